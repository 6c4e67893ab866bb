// The 3D part of the oracle: the tree of a 3D cloud, the power cell as a simple polytope, its measures against a constant
// density or a sum of Gaussians, and the sweep the Newton solver reads. Included by sdot_oracle.cpp inside `namespace orc`
// ( it uses its `select_median`, `preorder_of_heap`, `parallel_blocks`, `Facet`, `Events`, `Edges` ).
//
// Reference: the cell `cell/LocalN.h:320` ( simple polytope: every vertex on D planes with D neighbours; a cut keeps the
// inside vertices, makes one vertex per crossing edge and closes the new face by the planes the new vertices share ), the
// 3D card `gpu/Cell3D.cuh:417-730`, the walk and its pruning `cell/Providers.h:167-236`, `cell/Pruning.h:46-76` in three
// coordinates, the tree `AaBsp.py:264-680` / `bsp_build_level.h:84-318` with three axes, the sum of Gaussians
// `SumOfGaussians.cxx` ( densities `m / ( 2 pi s^2 )^( 3 / 2 ) exp( -|x - mu|^2 / ( 2 s^2 ) )` ).
//
// Integrals against a sum of Gaussians use fixed rules instead of the reference's adaptive `PointwiseDensity`
// ( `PointwiseDensity.h:31-134`, relative tolerance 1e-5 ): a cell's MASS is the flux through its faces of a field whose
// divergence is the density ( `flux_field` ), so that the masses of a diagram sum to the box's whatever the weights; a
// face integral uses the triangles from the face's area centroid and the 7-point degree-5 rule. The moments cut the cell in
// tetrahedra ( apex at the mean of its vertices, bases those triangles ) integrated by the 14-point degree-5 rule. Both rules are checked on every monomial of degree <= 5 by the tests.

struct Density3 {
    int           kind  = 0;             ///< 0 a constant, 1 a voxel image, 2 a sum of Gaussians
    double        value = 1;
    /// a voxel image ( `Image.cxx:106-232` in 3D ): voxel `( i, j, k )` covers `[ x0 + i hx, x0 + ( i + 1 ) hx ] x ...`, its value at
    /// `img[ ( k ny + j ) nx + i ]`; the cells are cut by every voxel they meet
    const double *img = nullptr;
    int           nx = 0, ny = 0, nz = 0;
    double        x0 = 0, y0 = 0, z0 = 0, hx = 1, hy = 1, hz = 1;
    int           ng = 0;
    const double *gauss = nullptr;       ///< [ ng ][ 5 ]: x, y, z, sigma, mass
    double        conv_s = 0;            ///< the continuation's width: sigma' = sqrt( sigma^2 + s^2 )
    double sigma( int q ) const { const double s = gauss[ 5 * q + 3 ]; return conv_s > 0 ? std::sqrt( s * s + conv_s * conv_s ) : s; }
    double at( double x, double y, double z ) const {
        if ( kind == 0 ) return value;
        double r = 0;
        for ( int q = 0; q < ng; ++q ) {
            const double *g = gauss + 5 * q;
            const double s = sigma( q ), dx = x - g[ 0 ], dy = y - g[ 1 ], dz = z - g[ 2 ];
            const double c = g[ 4 ] / ( 2.5066282746310002 * 2.5066282746310002 * 2.5066282746310002 * s * s * s );
            r += c * std::exp( -( dx * dx + dy * dy + dz * dz ) / ( 2 * s * s ) );
        }
        return r;
    }
};

/// `h( u ) = int_0^1 exp( -u t^2 ) t^2 dt`: its series below 1 ( no cancellation ), else `( sqrt( pi ) erf( sqrt u ) / 4 -
/// sqrt( u ) exp( -u ) / 2 ) / u^( 3 / 2 )`
static double radial_h( double u ) {
    if ( u < 1 ) {
        double term = 1, r = 1.0 / 3;
        for ( int k = 1; k < 40; ++k ) {
            term *= -u / k;
            const double c = term / ( 2 * k + 3 );
            r += c;
            if ( std::fabs( c ) < 1e-18 * r ) break;
        }
        return r;
    }
    const double su = std::sqrt( u );
    return ( 0.44311346272637900 * std::erf( su ) - 0.5 * su * std::exp( -u ) ) / ( u * su );
}

/// a field whose divergence is the density: per Gaussian `F( x ) = c ( x - mu ) h( |x - mu|^2 / ( 2 s^2 ) )`, `c` its peak
/// value ( the radial mass within `r`, over `r^2`, along `x - mu`; smooth at the centre ). The mass of a cell is then the
/// flux of `F` through its faces, and a face shared by two cells gives them opposite fluxes: the masses sum to the box's
/// whatever the weights ( the 3D counterpart of the 2D corner measures )
static void flux_field( const Density3 &dn, double x, double y, double z, double *F ) {
    F[ 0 ] = F[ 1 ] = F[ 2 ] = 0;
    for ( int q = 0; q < dn.ng; ++q ) {
        const double *g = dn.gauss + 5 * q;
        const double s = dn.sigma( q ), dx = x - g[ 0 ], dy = y - g[ 1 ], dz = z - g[ 2 ];
        const double c = g[ 4 ] / ( 2.5066282746310002 * 2.5066282746310002 * 2.5066282746310002 * s * s * s );
        const double f = c * radial_h( ( dx * dx + dy * dy + dz * dz ) / ( 2 * s * s ) );
        F[ 0 ] += f * dx; F[ 1 ] += f * dy; F[ 2 ] += f * dz;
    }
}

// ---- the quadrature rules ------------------------------------------------------------------------------------------

/// the 14-point degree-5 rule of the tetrahedron: barycentric points and weights summing to 1
struct TetRule { double b[ 14 ][ 4 ]; double w[ 14 ]; };
static const TetRule &tet_rule() {
    static const TetRule R = [] {
        TetRule r{};
        int k = 0;
        const double a1 = 0.0927352503108912264, w1 = 0.0734930431163619495;
        const double a2 = 0.3108859192633006098, w2 = 0.1126879257180158507;
        const double c = 0.0455037041256496494, w3 = 0.0425460207770814664;
        for ( double a : { a1, a2 } )
            for ( int j = 0; j < 4; ++j ) {
                for ( int i = 0; i < 4; ++i ) r.b[ k ][ i ] = i == j ? 1 - 3 * a : a;
                r.w[ k++ ] = a == a1 ? w1 : w2;
            }
        const int pairs[ 6 ][ 2 ] = { { 0, 1 }, { 0, 2 }, { 0, 3 }, { 1, 2 }, { 1, 3 }, { 2, 3 } };
        for ( const auto &p : pairs ) {
            for ( int i = 0; i < 4; ++i ) r.b[ k ][ i ] = ( i == p[ 0 ] || i == p[ 1 ] ) ? c : 0.5 - c;
            r.w[ k++ ] = w3;
        }
        return r;
    }();
    return R;
}

/// the 7-point degree-5 rule of the triangle ( Radon ): barycentric points and weights summing to 1
struct TriRule { double b[ 7 ][ 3 ]; double w[ 7 ]; };
static const TriRule &tri_rule() {
    static const TriRule R = [] {
        TriRule r{};
        const double s15 = std::sqrt( 15.0 );
        const double a1 = ( 6 - s15 ) / 21, w1 = ( 155 - s15 ) / 1200;
        const double a2 = ( 6 + s15 ) / 21, w2 = ( 155 + s15 ) / 1200;
        r.b[ 0 ][ 0 ] = r.b[ 0 ][ 1 ] = r.b[ 0 ][ 2 ] = 1.0 / 3;
        r.w[ 0 ] = 9.0 / 40;
        int k = 1;
        for ( double a : { a1, a2 } )
            for ( int j = 0; j < 3; ++j ) {
                for ( int i = 0; i < 3; ++i ) r.b[ k ][ i ] = i == j ? 1 - 2 * a : a;
                r.w[ k++ ] = a == a1 ? w1 : w2;
            }
        return r;
    }();
    return R;
}

// ---- the tree -----------------------------------------------------------------------------------------------------

struct Tree3 {
    I64                 n = 0, nb_nodes = 0;
    int                 depth = 0;
    std::vector<I64>    order, rank_of, beg, end;
    std::vector<double> box;             ///< [ N ][ 6 ]: lo x, y, z, hi x, y, z
    std::vector<double> wa, wb;          ///< [ N ][ 3 ], [ N ]
    std::vector<double> p;               ///< [ n ][ 3 ] positions in rank order
    std::vector<double> w;               ///< weights in rank order
};

/// the tree on `pos[ n ][ 3 ]`: median cut on the longest axis of the node's box ( the first on ties ), leaves of at most
/// `leaf` seeds, a perfect tree stored in preorder, seeds of a leaf in ascending user index
void build_tree3( Tree3 &t, I64 n, const double *pos, I64 leaf ) {
    const int depth = max_depth_for( n, leaf );
    const I64 nb_nodes = ( I64( 1 ) << depth ) - 1;
    t.n = n; t.depth = depth; t.nb_nodes = nb_nodes;
    std::vector<double> P[ 3 ], P2[ 3 ];
    for ( int d = 0; d < 3; ++d ) { P[ d ].resize( sz( n ) ); P2[ d ].resize( sz( n ) ); }
    std::vector<I64> ord( sz( n ) ), ord2( sz( n ) ), perm( sz( n ) );
    for ( I64 i = 0; i < n; ++i ) { for ( int d = 0; d < 3; ++d ) P[ d ][ sz( i ) ] = pos[ 3 * i + d ]; ord[ sz( i ) ] = i; }
    std::vector<I64> hbeg( sz( nb_nodes ) ), hend( sz( nb_nodes ) );
    std::vector<double> hbox( sz( nb_nodes ) * 6, 0.0 );
    std::vector<I64> beg{ 0 }, end{ n };
    for ( int level = 0; level < depth; ++level ) {
        const I64 first = ( I64( 1 ) << level ) - 1;
        std::vector<I64> mid( end );
        for ( size_t j = 0; j < beg.size(); ++j ) {
            const I64 b = beg[ j ], e = end[ j ], node = first + I64( j );
            hbeg[ sz( node ) ] = b;
            hend[ sz( node ) ] = e;
            if ( e <= b ) continue;
            double lo[ 3 ], hi[ 3 ];
            for ( int d = 0; d < 3; ++d ) lo[ d ] = hi[ d ] = P[ d ][ sz( b ) ];
            for ( I64 k = b + 1; k < e; ++k )
                for ( int d = 0; d < 3; ++d ) { lo[ d ] = std::min( lo[ d ], P[ d ][ sz( k ) ] ); hi[ d ] = std::max( hi[ d ], P[ d ][ sz( k ) ] ); }
            double *bx = &hbox[ sz( node ) * 6 ];
            for ( int d = 0; d < 3; ++d ) { bx[ d ] = lo[ d ]; bx[ 3 + d ] = hi[ d ]; }
            int ax = 0;
            for ( int d = 1; d < 3; ++d ) if ( hi[ d ] - lo[ d ] > hi[ ax ] - lo[ ax ] ) ax = d;
            for ( I64 k = b; k < e; ++k ) perm[ sz( k ) ] = k;
            I64 m = e;
            if ( e - b > leaf && hi[ ax ] > lo[ ax ] ) {
                m = b + ( e - b ) / 2;
                select_median( perm, P[ ax ], b, e, m );
            }
            mid[ j ] = m;
            for ( I64 k = b; k < e; ++k ) {
                const I64 s = perm[ sz( k ) ];
                for ( int d = 0; d < 3; ++d ) P2[ d ][ sz( k ) ] = P[ d ][ sz( s ) ];
                ord2[ sz( k ) ] = ord[ sz( s ) ];
            }
        }
        for ( int d = 0; d < 3; ++d ) P[ d ].swap( P2[ d ] );
        ord.swap( ord2 );
        if ( level + 1 < depth ) {
            std::vector<I64> nb_( 2 * beg.size() ), ne_( 2 * beg.size() );
            for ( size_t j = 0; j < beg.size(); ++j ) {
                nb_[ 2 * j ] = beg[ j ]; ne_[ 2 * j ] = mid[ j ];
                nb_[ 2 * j + 1 ] = mid[ j ]; ne_[ 2 * j + 1 ] = end[ j ];
            }
            beg.swap( nb_ ); end.swap( ne_ );
        }
    }
    {
        const I64 first = ( I64( 1 ) << ( depth - 1 ) ) - 1;
        std::vector<I64> idx;
        for ( I64 h = first; h < nb_nodes; ++h ) {
            const I64 b = hbeg[ sz( h ) ], e = hend[ sz( h ) ];
            if ( e - b < 2 ) continue;
            idx.resize( sz( e - b ) );
            for ( I64 k = b; k < e; ++k ) idx[ sz( k - b ) ] = k;
            std::sort( idx.begin(), idx.end(), [&]( I64 u, I64 v ) { return ord[ sz( u ) ] < ord[ sz( v ) ]; } );
            std::vector<double> tp[ 3 ];
            std::vector<I64> to( idx.size() );
            for ( int d = 0; d < 3; ++d ) tp[ d ].resize( idx.size() );
            for ( size_t q = 0; q < idx.size(); ++q ) {
                for ( int d = 0; d < 3; ++d ) tp[ d ][ q ] = P[ d ][ sz( idx[ q ] ) ];
                to[ q ] = ord[ sz( idx[ q ] ) ];
            }
            for ( size_t q = 0; q < idx.size(); ++q ) {
                for ( int d = 0; d < 3; ++d ) P[ d ][ sz( b ) + q ] = tp[ d ][ q ];
                ord[ sz( b ) + q ] = to[ q ];
            }
        }
    }
    const std::vector<I64> pre = preorder_of_heap( depth );
    t.beg.assign( sz( nb_nodes ), 0 );
    t.end.assign( sz( nb_nodes ), 0 );
    t.box.assign( sz( nb_nodes ) * 6, 0.0 );
    for ( I64 h = 0; h < nb_nodes; ++h ) {
        const I64 q = pre[ sz( h ) ];
        t.beg[ sz( q ) ] = hbeg[ sz( h ) ];
        t.end[ sz( q ) ] = hend[ sz( h ) ];
        for ( int c = 0; c < 6; ++c ) t.box[ sz( q ) * 6 + sz( c ) ] = hbox[ sz( h ) * 6 + sz( c ) ];
    }
    t.order = ord;
    t.rank_of.assign( sz( n ), 0 );
    for ( I64 k = 0; k < n; ++k ) t.rank_of[ sz( ord[ sz( k ) ] ) ] = k;
    t.p.assign( sz( n ) * 3, 0.0 );
    for ( I64 k = 0; k < n; ++k ) for ( int d = 0; d < 3; ++d ) t.p[ sz( 3 * k + d ) ] = P[ d ][ sz( k ) ];
    t.w.assign( sz( n ), 0.0 );
    t.wa.assign( sz( nb_nodes ) * 3, 0.0 );
    t.wb.assign( sz( nb_nodes ), 0.0 );
}

/// the affine majorant `w <= a . y + b` of the rank slice `[ b, e )` in 3D ( `bsp_build_level.h:84-209`: the least-squares
/// slope kept when the residuals' spread is under `0.85 sqrt( 1 - 3 / ( m - 1 ) )` times the weights' and the slope is well
/// behaved on the node; `b` raised to majorize, plus `1e-6 ( |b| + spread + max |a . y| )` )
static void weight_majorant3( const double *p, const double *w, I64 b, I64 e, double *wa, double &wb ) {
    const I64 m = e - b;
    double wmin = w[ b ], wmax = w[ b ], wsum = 0, ps[ 3 ] = { 0, 0, 0 }, plo[ 3 ], phi[ 3 ];
    for ( int d = 0; d < 3; ++d ) plo[ d ] = phi[ d ] = p[ 3 * b + d ];
    for ( I64 k = b; k < e; ++k ) {
        wmin = std::min( wmin, w[ k ] ); wmax = std::max( wmax, w[ k ] ); wsum += w[ k ];
        for ( int d = 0; d < 3; ++d ) {
            ps[ d ] += p[ 3 * k + d ];
            plo[ d ] = std::min( plo[ d ], p[ 3 * k + d ] ); phi[ d ] = std::max( phi[ d ], p[ 3 * k + d ] );
        }
    }
    const double spread = wmax - wmin;
    double a[ 3 ] = { 0, 0, 0 };
    if ( m >= 2 * ( 3 + 1 ) && spread > 0 ) {
        const double inv = 1.0 / double( m );
        const double pm[ 3 ] = { ps[ 0 ] * inv, ps[ 1 ] * inv, ps[ 2 ] * inv }, wm = wsum * inv;
        double A[ 3 ][ 4 ] = {};
        for ( I64 k = b; k < e; ++k ) {
            const double dw = w[ k ] - wm;
            const double q[ 3 ] = { p[ 3 * k ] - pm[ 0 ], p[ 3 * k + 1 ] - pm[ 1 ], p[ 3 * k + 2 ] - pm[ 2 ] };
            for ( int i = 0; i < 3; ++i ) {
                for ( int j = 0; j < 3; ++j ) A[ i ][ j ] += q[ i ] * q[ j ];
                A[ i ][ 3 ] += q[ i ] * dw;
            }
        }
        bool ok = true;
        for ( int c = 0; c < 3 && ok; ++c ) {
            int piv = c;
            for ( int i = c + 1; i < 3; ++i ) if ( std::fabs( A[ i ][ c ] ) > std::fabs( A[ piv ][ c ] ) ) piv = i;
            if ( ! ( std::fabs( A[ piv ][ c ] ) > 0 ) ) { ok = false; break; }
            if ( piv != c ) for ( int j = c; j <= 3; ++j ) std::swap( A[ c ][ j ], A[ piv ][ j ] );
            for ( int i = c + 1; i < 3; ++i ) {
                const double f = A[ i ][ c ] / A[ c ][ c ];
                for ( int j = c; j <= 3; ++j ) A[ i ][ j ] -= f * A[ c ][ j ];
            }
        }
        if ( ok ) {
            double fit[ 3 ];
            for ( int i = 2; i >= 0; --i ) {
                double s = A[ i ][ 3 ];
                for ( int j = i + 1; j < 3; ++j ) s -= A[ i ][ j ] * fit[ j ];
                fit[ i ] = s / A[ i ][ i ];
            }
            double rmin = 0, rmax = 0;
            for ( I64 k = b; k < e; ++k ) {
                const double r = w[ k ] - fit[ 0 ] * p[ 3 * k ] - fit[ 1 ] * p[ 3 * k + 1 ] - fit[ 2 ] * p[ 3 * k + 2 ];
                if ( k == b ) { rmin = r; rmax = r; }
                else { rmin = std::min( rmin, r ); rmax = std::max( rmax, r ); }
            }
            const double u = 1.0 - 3.0 / double( m - 1 );
            const double by_chance = std::sqrt( u > 0 ? u : 0.0 );
            bool well_behaved = true;
            for ( int d = 0; d < 3; ++d ) {
                const double reach = std::max( std::fabs( plo[ d ] ), std::fabs( phi[ d ] ) );
                if ( std::fabs( fit[ d ] ) * ( phi[ d ] - plo[ d ] ) > 8 * spread || std::fabs( fit[ d ] ) * reach > 100 * spread ) well_behaved = false;
            }
            if ( well_behaved && rmax - rmin < 0.85 * by_chance * spread ) for ( int d = 0; d < 3; ++d ) a[ d ] = fit[ d ];
        }
    }
    double bb = 0, amax = 0;
    for ( I64 k = b; k < e; ++k ) {
        const double ay = a[ 0 ] * p[ 3 * k ] + a[ 1 ] * p[ 3 * k + 1 ] + a[ 2 ] * p[ 3 * k + 2 ];
        const double v = w[ k ] - ay;
        bb = k == b ? v : std::max( bb, v );
        amax = std::max( amax, std::fabs( ay ) );
    }
    for ( int d = 0; d < 3; ++d ) wa[ d ] = a[ d ];
    wb = bb + 1e-6 * ( std::fabs( bb ) + spread + amax );
}

void set_weights3( Tree3 &t, const double *w_user ) {
    for ( I64 k = 0; k < t.n; ++k ) t.w[ sz( k ) ] = w_user[ t.order[ sz( k ) ] ];
    parallel_blocks( t.nb_nodes, 64, [&]( int, I64 b, I64 e ) {
        for ( I64 m = b; m < e; ++m ) {
            const I64 sb = t.beg[ sz( m ) ], se = t.end[ sz( m ) ];
            if ( se <= sb ) { for ( int d = 0; d < 3; ++d ) t.wa[ sz( 3 * m + d ) ] = 0; t.wb[ sz( m ) ] = 0; continue; }
            weight_majorant3( t.p.data(), t.w.data(), sb, se, &t.wa[ sz( 3 * m ) ], t.wb[ sz( m ) ] );
        }
    } );
}

// ---- the polytope -------------------------------------------------------------------------------------------------

/// a plane of the cell, `n . v <= off` in the seed's frame; `id` a seed rank, or `-1 - f` face `f` of the domain box
/// ( 0 x lo, 1 x hi, 2 y lo, 3 y hi, 4 z lo, 5 z hi )
struct Face3 { double nx, ny, nz, off; int id; };

/// a simple convex polytope in the seed's frame: vertex `v` lies on the faces `pl[ 3 v + j ]`, and `nb[ 3 v + j ]` is the
/// vertex at the other end of its edge that LEAVES face `pl[ 3 v + j ]` ( the edge along the other two )
struct Poly3 {
    int                 nv = 0;
    std::vector<double> x, y, z, s;
    std::vector<int>    pl, nb;
    std::vector<Face3>  f;
    bool                broken = false;      ///< a cut whose new face could not be closed ( inconsistent signs )
};

static void init_box3( Poly3 &c, const double *lo, const double *hi, const double *p0 ) {
    c.nv = 8;
    c.x.assign( 8, 0 ); c.y.assign( 8, 0 ); c.z.assign( 8, 0 ); c.s.assign( 8, 0 );
    c.pl.assign( 24, 0 ); c.nb.assign( 24, 0 );
    c.f.clear();
    c.broken = false;
    for ( int d = 0; d < 3; ++d ) {
        Face3 a{ 0, 0, 0, p0[ d ] - lo[ d ], -1 - 2 * d }, b{ 0, 0, 0, hi[ d ] - p0[ d ], -2 - 2 * d };
        ( d == 0 ? a.nx : d == 1 ? a.ny : a.nz ) = -1;
        ( d == 0 ? b.nx : d == 1 ? b.ny : b.nz ) = 1;
        c.f.push_back( a );
        c.f.push_back( b );
    }
    for ( int v = 0; v < 8; ++v ) {
        c.x[ sz( v ) ] = ( v & 1 ? hi[ 0 ] : lo[ 0 ] ) - p0[ 0 ];
        c.y[ sz( v ) ] = ( v & 2 ? hi[ 1 ] : lo[ 1 ] ) - p0[ 1 ];
        c.z[ sz( v ) ] = ( v & 4 ? hi[ 2 ] : lo[ 2 ] ) - p0[ 2 ];
        for ( int d = 0; d < 3; ++d ) {
            c.pl[ sz( 3 * v + d ) ] = 2 * d + ( ( v >> d ) & 1 );
            c.nb[ sz( 3 * v + d ) ] = v ^ ( 1 << d );
        }
    }
}

/// the cut by `n . v <= off` ( `cell/LocalN.h:320-420` ): a vertex is outside when `s > 0`; a crossing edge `( a in, b out )`
/// gets a vertex interpolated from `a`, on the two faces the edge lies on and the new face; two new vertices are joined
/// when they share one of those faces
static int cut3( Poly3 &c, double nx, double ny, double nz, double off, int id, Events *ev ) {
    const int nv = c.nv;
    int nb_out = 0;
    double vmax = 0;
    for ( int v = 0; v < nv; ++v ) {
        c.s[ sz( v ) ] = std::fma( nx, c.x[ sz( v ) ], std::fma( ny, c.y[ sz( v ) ], std::fma( nz, c.z[ sz( v ) ], -off ) ) );
        vmax = std::max( vmax, std::max( std::fabs( c.x[ sz( v ) ] ), std::max( std::fabs( c.y[ sz( v ) ] ), std::fabs( c.z[ sz( v ) ] ) ) ) );
    }
    // a vertex on the plane up to the rounding of `s` stays inside: where more than three planes meet ( a grid's corners ),
    // the signs of near-zero `s` would otherwise split a face's outside vertices in two runs
    const double tie = 1e-13 * ( ( std::fabs( nx ) + std::fabs( ny ) + std::fabs( nz ) ) * vmax + std::fabs( off ) );
    for ( int v = 0; v < nv; ++v ) {
        if ( ! ( c.s[ sz( v ) ] > tie ) ) c.s[ sz( v ) ] = std::min( c.s[ sz( v ) ], 0.0 );
        nb_out += c.s[ sz( v ) ] > 0;
    }
    if ( ev ) ev->side_tests += nv;
    if ( nb_out == 0 ) return UNCHANGED;
    if ( nb_out == nv ) { c.nv = 0; return EMPTY; }
    const int fn = int( c.f.size() );
    c.f.push_back( Face3{ nx, ny, nz, off, id } );
    std::vector<int> idx( sz( nv ), -1 );
    int nin = 0;
    for ( int v = 0; v < nv; ++v ) if ( ! ( c.s[ sz( v ) ] > 0 ) ) idx[ sz( v ) ] = nin++;
    // the new vertices, in the order of ( outside vertex, edge )
    struct New { double x, y, z; int p1, p2, a; };
    std::vector<New> nw;
    std::vector<int> nnb( sz( 3 * nin ) ), npl( sz( 3 * nin ) );
    for ( int v = 0; v < nv; ++v ) {
        if ( idx[ sz( v ) ] < 0 ) continue;
        for ( int j = 0; j < 3; ++j ) {
            npl[ sz( 3 * idx[ sz( v ) ] + j ) ] = c.pl[ sz( 3 * v + j ) ];
            const int u = c.nb[ sz( 3 * v + j ) ];
            nnb[ sz( 3 * idx[ sz( v ) ] + j ) ] = idx[ sz( u ) ];   // ( an outside neighbour is replaced below )
        }
    }
    for ( int b = 0; b < nv; ++b ) {
        if ( idx[ sz( b ) ] >= 0 ) continue;
        for ( int k = 0; k < 3; ++k ) {
            const int a = c.nb[ sz( 3 * b + k ) ];
            if ( idx[ sz( a ) ] < 0 ) continue;
            const double sa = c.s[ sz( a ) ], sb = c.s[ sz( b ) ], t = sa / ( sa - sb );
            New q;
            q.x = std::fma( c.x[ sz( b ) ] - c.x[ sz( a ) ], t, c.x[ sz( a ) ] );
            q.y = std::fma( c.y[ sz( b ) ] - c.y[ sz( a ) ], t, c.y[ sz( a ) ] );
            q.z = std::fma( c.z[ sz( b ) ] - c.z[ sz( a ) ], t, c.z[ sz( a ) ] );
            q.p1 = c.pl[ sz( 3 * b + ( k + 1 ) % 3 ) ];
            q.p2 = c.pl[ sz( 3 * b + ( k + 2 ) % 3 ) ];
            q.a = idx[ sz( a ) ];
            for ( int j = 0; j < 3; ++j )
                if ( c.nb[ sz( 3 * a + j ) ] == b ) nnb[ sz( 3 * q.a + j ) ] = nin + int( nw.size() );
            nw.push_back( q );
        }
    }
    const int T = int( nw.size() ), m = nin + T;
    std::vector<double> X( sz( m ) ), Y( sz( m ) ), Z( sz( m ) );
    for ( int v = 0; v < nv; ++v )
        if ( idx[ sz( v ) ] >= 0 ) { X[ sz( idx[ sz( v ) ] ) ] = c.x[ sz( v ) ]; Y[ sz( idx[ sz( v ) ] ) ] = c.y[ sz( v ) ]; Z[ sz( idx[ sz( v ) ] ) ] = c.z[ sz( v ) ]; }
    npl.resize( sz( 3 * m ) );
    nnb.resize( sz( 3 * m ) );
    for ( int u = 0; u < T; ++u ) {
        const New &q = nw[ sz( u ) ];
        const int g = nin + u;
        X[ sz( g ) ] = q.x; Y[ sz( g ) ] = q.y; Z[ sz( g ) ] = q.z;
        npl[ sz( 3 * g ) ] = fn; npl[ sz( 3 * g + 1 ) ] = q.p1; npl[ sz( 3 * g + 2 ) ] = q.p2;
        nnb[ sz( 3 * g ) ] = q.a;
        // leaving p1 ( along the new face and p2 ): the other new vertex on p2; leaving p2: the other on p1
        int m1 = -1, m2 = -1, c1 = 0, c2 = 0;
        for ( int o = 0; o < T; ++o ) {
            if ( o == u ) continue;
            const New &r = nw[ sz( o ) ];
            if ( r.p1 == q.p2 || r.p2 == q.p2 ) { m1 = nin + o; ++c1; }
            if ( r.p1 == q.p1 || r.p2 == q.p1 ) { m2 = nin + o; ++c2; }
        }
        if ( c1 != 1 || c2 != 1 ) c.broken = true;
        nnb[ sz( 3 * g + 1 ) ] = m1;
        nnb[ sz( 3 * g + 2 ) ] = m2;
    }
    c.nv = m;
    c.x.swap( X ); c.y.swap( Y ); c.z.swap( Z );
    c.s.assign( sz( m ), 0.0 );
    c.pl.swap( npl );
    c.nb.swap( nnb );
    if ( ev ) { ev->new_vertices += T; ev->cuts += 1; }
    return CUT;
}

/// the polygon of face `f` ( vertex indices in a cyclic order; empty if the face has no vertex )
static void face_loop( const Poly3 &c, int f, std::vector<int> &loop ) {
    loop.clear();
    int v0 = -1;
    for ( int v = 0; v < c.nv && v0 < 0; ++v )
        for ( int j = 0; j < 3; ++j ) if ( c.pl[ sz( 3 * v + j ) ] == f ) { v0 = v; break; }
    if ( v0 < 0 ) return;
    int prev = -1, cur = v0;
    for ( int guard = 0; guard <= c.nv; ++guard ) {
        loop.push_back( cur );
        int j = 0;
        while ( j < 3 && c.pl[ sz( 3 * cur + j ) ] != f ) ++j;
        if ( j == 3 ) { loop.clear(); return; }
        const int a = c.nb[ sz( 3 * cur + ( j + 1 ) % 3 ) ], b = c.nb[ sz( 3 * cur + ( j + 2 ) % 3 ) ];
        const int next = a != prev ? a : b;
        prev = cur;
        cur = next;
        if ( cur == v0 ) return;
        if ( cur < 0 ) { loop.clear(); return; }
    }
    loop.clear();
}

/// the area centroid of a face polygon ( its loop of vertices ), from the triangles of its mean vertex: the apex of the
/// face's triangles for the rules. The mean of the vertices would jump when a vertex appears or vanishes with an edge of
/// zero length, and the rules' error with it ( ~`( h / sigma )^5` of a cell's mass for a Gaussian, `h` the cell's size ):
/// a jump of the masses in the weights, which stalls a Newton solve asked for less than that error; the centroid moves with
/// the vertices
static void face_centroid( const Poly3 &c, const std::vector<int> &loop, double *G ) {
    const size_t nl = loop.size();
    double gx = 0, gy = 0, gz = 0;
    for ( int v : loop ) { gx += c.x[ sz( v ) ]; gy += c.y[ sz( v ) ]; gz += c.z[ sz( v ) ]; }
    gx /= double( nl ); gy /= double( nl ); gz /= double( nl );
    double sa = 0, sx = 0, sy = 0, sw = 0;
    for ( size_t i = 0; i < nl; ++i ) {
        const int v1 = loop[ i ], v2 = loop[ ( i + 1 ) % nl ];
        const double ux = c.x[ sz( v1 ) ] - gx, uy = c.y[ sz( v1 ) ] - gy, uz = c.z[ sz( v1 ) ] - gz;
        const double vx = c.x[ sz( v2 ) ] - gx, vy = c.y[ sz( v2 ) ] - gy, vz = c.z[ sz( v2 ) ] - gz;
        const double cx = uy * vz - uz * vy, cy = uz * vx - ux * vz, cz = ux * vy - uy * vx;
        const double a = 0.5 * std::sqrt( cx * cx + cy * cy + cz * cz );
        sa += a; sx += a * ( ux + vx ); sy += a * ( uy + vy ); sw += a * ( uz + vz );
    }
    if ( sa > 0 ) { gx += sx / ( 3 * sa ); gy += sy / ( 3 * sa ); gz += sw / ( 3 * sa ); }
    G[ 0 ] = gx; G[ 1 ] = gy; G[ 2 ] = gz;
}

/// for every face: `f( face, loop )`; then its tetrahedra ( apex `A` the mean of the cell's vertices, base the triangle
/// of the face's area centroid and one of its edges ) with their volumes: `t( A, B, C, D, vol )`. Both in the seed's frame.
template<class FF, class TF>
static void for_each_face_and_tet( const Poly3 &c, FF &&ff, TF &&tf ) {
    double ax = 0, ay = 0, az = 0;
    for ( int v = 0; v < c.nv; ++v ) { ax += c.x[ sz( v ) ]; ay += c.y[ sz( v ) ]; az += c.z[ sz( v ) ]; }
    ax /= c.nv; ay /= c.nv; az /= c.nv;
    const double A[ 3 ] = { ax, ay, az };
    std::vector<int> loop;
    for ( int f = 0; f < int( c.f.size() ); ++f ) {
        face_loop( c, f, loop );
        if ( loop.size() < 3 ) continue;
        ff( f, loop );
        const size_t nl = loop.size();
        double B[ 3 ];
        face_centroid( c, loop, B );
        for ( size_t i = 0; i < nl; ++i ) {
            const int v1 = loop[ i ], v2 = loop[ ( i + 1 ) % nl ];
            const double C[ 3 ] = { c.x[ sz( v1 ) ], c.y[ sz( v1 ) ], c.z[ sz( v1 ) ] };
            const double D[ 3 ] = { c.x[ sz( v2 ) ], c.y[ sz( v2 ) ], c.z[ sz( v2 ) ] };
            const double u[ 3 ] = { B[ 0 ] - A[ 0 ], B[ 1 ] - A[ 1 ], B[ 2 ] - A[ 2 ] };
            const double v[ 3 ] = { C[ 0 ] - A[ 0 ], C[ 1 ] - A[ 1 ], C[ 2 ] - A[ 2 ] };
            const double w[ 3 ] = { D[ 0 ] - A[ 0 ], D[ 1 ] - A[ 1 ], D[ 2 ] - A[ 2 ] };
            const double det = u[ 0 ] * ( v[ 1 ] * w[ 2 ] - v[ 2 ] * w[ 1 ] ) - u[ 1 ] * ( v[ 0 ] * w[ 2 ] - v[ 2 ] * w[ 0 ] ) + u[ 2 ] * ( v[ 0 ] * w[ 1 ] - v[ 1 ] * w[ 0 ] );
            tf( A, B, C, D, std::fabs( det ) / 6 );
        }
    }
}

/// `int_T rho`, `int_T x rho`, `int_T |x|^2 rho` over a tetrahedron ( seed frame, the density read at `p0 + x` )
static void tet_moments( const Density3 &dn, const double *p0, const double *A, const double *B, const double *C, const double *D, double vol,
                         double &m, double *mx, double &m2 ) {
    if ( dn.kind == 0 ) {                                // exact: the centroid and `V / 20 ( sum |v_i|^2 + |sum v_i|^2 )`
        const double w = dn.value * vol;
        double s[ 3 ], sq = 0, ss = 0;
        for ( int d = 0; d < 3; ++d ) {
            s[ d ] = A[ d ] + B[ d ] + C[ d ] + D[ d ];
            sq += A[ d ] * A[ d ] + B[ d ] * B[ d ] + C[ d ] * C[ d ] + D[ d ] * D[ d ];
            ss += s[ d ] * s[ d ];
            mx[ d ] += w * s[ d ] / 4;
        }
        m += w;
        m2 += w * ( sq + ss ) / 20;
        return;
    }
    const TetRule &R = tet_rule();
    for ( int q = 0; q < 14; ++q ) {
        double x[ 3 ];
        for ( int d = 0; d < 3; ++d ) x[ d ] = R.b[ q ][ 0 ] * A[ d ] + R.b[ q ][ 1 ] * B[ d ] + R.b[ q ][ 2 ] * C[ d ] + R.b[ q ][ 3 ] * D[ d ];
        const double w = R.w[ q ] * vol * dn.at( p0[ 0 ] + x[ 0 ], p0[ 1 ] + x[ 1 ], p0[ 2 ] + x[ 2 ] );
        m += w;
        for ( int d = 0; d < 3; ++d ) mx[ d ] += w * x[ d ];
        m2 += w * ( x[ 0 ] * x[ 0 ] + x[ 1 ] * x[ 1 ] + x[ 2 ] * x[ 2 ] );
    }
}

/// the box's faces carry the flux out of the domain: their triangles are integrated on `BOX_SUB x BOX_SUB` sub-triangles,
/// so that the masses' total ( the box's flux, split among the cells the box faces touch ) hardly depends on the weights
constexpr int BOX_SUB = 4;

/// `int_face rho` over a face polygon ( seed frame ), or with `flux` the flux `int_face F . n` of the field above through
/// it ( `n` the face's outward unit normal ): triangles from the polygon's area centroid ( `face_centroid` ), so that the two
/// cells of a face integrate it alike whatever vertex their loops start from
static double face_integral( const Density3 &dn, const double *p0, const Poly3 &c, const std::vector<int> &loop, const double *flux = nullptr,
                             int sub = 1 ) {
    const size_t nl = loop.size();
    double G[ 3 ];
    face_centroid( c, loop, G );
    const double gx = G[ 0 ], gy = G[ 1 ], gz = G[ 2 ];
    double r = 0;
    for ( size_t i = 0; i < nl; ++i ) {
        const int v1 = loop[ i ], v2 = loop[ ( i + 1 ) % nl ];
        const double ux = c.x[ sz( v1 ) ] - gx, uy = c.y[ sz( v1 ) ] - gy, uz = c.z[ sz( v1 ) ] - gz;
        const double vx = c.x[ sz( v2 ) ] - gx, vy = c.y[ sz( v2 ) ] - gy, vz = c.z[ sz( v2 ) ] - gz;
        const double cx = uy * vz - uz * vy, cy = uz * vx - ux * vz, cz = ux * vy - uy * vx;
        const double ar = 0.5 * std::sqrt( cx * cx + cy * cy + cz * cz );
        if ( dn.kind == 0 && ! flux ) { r += dn.value * ar; continue; }
        // the triangle on an `m x m` grid of sub-triangles ( up `( i, j )`, `( i + 1, j )`, `( i, j + 1 )`, down
        // `( i + 1, j )`, `( i + 1, j + 1 )`, `( i, j + 1 )` ), each by the 7-point rule
        const TriRule &R = tri_rule();
        const double A[ 3 ] = { gx, gy, gz }, B[ 3 ] = { c.x[ sz( v1 ) ], c.y[ sz( v1 ) ], c.z[ sz( v1 ) ] }, C[ 3 ] = { c.x[ sz( v2 ) ], c.y[ sz( v2 ) ], c.z[ sz( v2 ) ] };
        auto at = [&]( int i, int j, double *P ) { for ( int d = 0; d < 3; ++d ) P[ d ] = A[ d ] + ( B[ d ] - A[ d ] ) * i / sub + ( C[ d ] - A[ d ] ) * j / sub; };
        const double sar = ar / ( sub * sub );
        for ( int i = 0; i < sub; ++i )
            for ( int j = 0; i + j < sub; ++j )
                for ( int down = 0; down < ( i + j + 1 < sub ? 2 : 1 ); ++down ) {
                    double P0[ 3 ], P1[ 3 ], P2[ 3 ];
                    if ( ! down ) { at( i, j, P0 ); at( i + 1, j, P1 ); at( i, j + 1, P2 ); }
                    else          { at( i + 1, j, P0 ); at( i + 1, j + 1, P1 ); at( i, j + 1, P2 ); }
                    for ( int q = 0; q < 7; ++q ) {
                        const double x = R.b[ q ][ 0 ] * P0[ 0 ] + R.b[ q ][ 1 ] * P1[ 0 ] + R.b[ q ][ 2 ] * P2[ 0 ];
                        const double y = R.b[ q ][ 0 ] * P0[ 1 ] + R.b[ q ][ 1 ] * P1[ 1 ] + R.b[ q ][ 2 ] * P2[ 1 ];
                        const double z = R.b[ q ][ 0 ] * P0[ 2 ] + R.b[ q ][ 1 ] * P1[ 2 ] + R.b[ q ][ 2 ] * P2[ 2 ];
                        if ( flux ) {
                            double F[ 3 ];
                            flux_field( dn, p0[ 0 ] + x, p0[ 1 ] + y, p0[ 2 ] + z, F );
                            r += R.w[ q ] * sar * ( F[ 0 ] * flux[ 0 ] + F[ 1 ] * flux[ 1 ] + F[ 2 ] * flux[ 2 ] );
                        } else
                            r += R.w[ q ] * sar * dn.at( p0[ 0 ] + x, p0[ 1 ] + y, p0[ 2 ] + z );
                    }
                }
    }
    return r;
}

// ---- a voxel image: the cell cut by every voxel it meets ------------------------------------------------------------

/// the voxels of the cell ( seed frame, the seed at `p0` ), each cut by its six planes: `f( value, piece )` per non-empty piece.
/// The image is zero off its grid: what lies outside it is in no piece
template<class F>
static void for_each_voxel_piece( const Density3 &dn, const double *p0, const Poly3 &c, F &&f ) {
    double lo[ 3 ] = { 1e300, 1e300, 1e300 }, hi[ 3 ] = { -1e300, -1e300, -1e300 };
    for ( int v = 0; v < c.nv; ++v ) {
        const double q[ 3 ] = { c.x[ sz( v ) ], c.y[ sz( v ) ], c.z[ sz( v ) ] };
        for ( int d = 0; d < 3; ++d ) { lo[ d ] = std::min( lo[ d ], q[ d ] ); hi[ d ] = std::max( hi[ d ], q[ d ] ); }
    }
    const double o[ 3 ] = { dn.x0 - p0[ 0 ], dn.y0 - p0[ 1 ], dn.z0 - p0[ 2 ] }, h[ 3 ] = { dn.hx, dn.hy, dn.hz };
    const int nn[ 3 ] = { dn.nx, dn.ny, dn.nz };
    int a[ 3 ], b[ 3 ];
    for ( int d = 0; d < 3; ++d ) {
        if ( hi[ d ] < o[ d ] || lo[ d ] > o[ d ] + nn[ d ] * h[ d ] ) return;      // ( the image is zero outside its grid )
        auto vox = [&]( double x ) { const double t = std::floor( ( x - o[ d ] ) / h[ d ] ); return t < 0 ? 0 : ( t >= nn[ d ] - 1 ? nn[ d ] - 1 : int( t ) ); };
        a[ d ] = vox( lo[ d ] ); b[ d ] = vox( hi[ d ] );
    }
    Poly3 p;
    for ( int k = a[ 2 ]; k <= b[ 2 ]; ++k )
        for ( int j = a[ 1 ]; j <= b[ 1 ]; ++j )
            for ( int i = a[ 0 ]; i <= b[ 0 ]; ++i ) {
                p = c;
                const int idx[ 3 ] = { i, j, k };
                bool empty = false;
                for ( int d = 0; d < 3 && ! empty; ++d ) {            // ( the voxel's two planes along each axis )
                    double n[ 3 ] = { 0, 0, 0 };
                    n[ d ] = -1;
                    empty = cut3( p, n[ 0 ], n[ 1 ], n[ 2 ], -( o[ d ] + idx[ d ] * h[ d ] ), -100 - 2 * d, nullptr ) == EMPTY;
                    if ( ! empty ) {
                        n[ d ] = 1;
                        empty = cut3( p, n[ 0 ], n[ 1 ], n[ 2 ], o[ d ] + ( idx[ d ] + 1 ) * h[ d ], -101 - 2 * d, nullptr ) == EMPTY;
                    }
                }
                if ( ! empty && p.nv >= 4 ) f( dn.img[ ( size_t( k ) * size_t( dn.ny ) + size_t( j ) ) * size_t( dn.nx ) + size_t( i ) ], p );
            }
}

// ---- the walk -----------------------------------------------------------------------------------------------------

static bool can_cut3( const Tree3 &t, I64 m, const Poly3 &c, const double *p0, double w0, Events *ev ) {
    const double *bx = &t.box[ sz( m ) * 6 ];
    const double *a = &t.wa[ sz( 3 * m ) ];
    const double b = t.wb[ sz( m ) ];
    if ( ev ) ev->node_visits += 1;
    for ( int i = 0; i < c.nv; ++i ) {
        if ( ev ) ev->prune_vertex_tests += 1;
        const double v[ 3 ] = { c.x[ sz( i ) ], c.y[ sz( i ) ], c.z[ sz( i ) ] };
        double s = w0 - b;
        for ( int d = 0; d < 3; ++d ) {
            const double X = v[ d ] + p0[ d ];
            double y = X + a[ d ] / 2;
            y = y < bx[ d ] ? bx[ d ] : ( y > bx[ 3 + d ] ? bx[ 3 + d ] : y );
            const double u = y - X;
            s += u * u - v[ d ] * v[ d ] - a[ d ] * y;
        }
        if ( s <= 0 ) return true;
    }
    return false;
}

static double proximity3( const Tree3 &t, I64 m, const double *p0 ) {
    const double *bx = &t.box[ sz( m ) * 6 ];
    double r = 0;
    for ( int d = 0; d < 3; ++d ) {
        double e = std::max( bx[ d ] - p0[ d ], p0[ d ] - bx[ 3 + d ] );
        e = e > 0 ? e : 0;
        r += e * e;
    }
    return r;
}

/// the cell of rank `k0`: the domain box cut by the planes of the seeds the tree proposes, nearest node first
bool make_cell3( const Tree3 &t, I64 k0, const double *lo, const double *hi, Poly3 &c, Events *ev ) {
    const double *p0 = &t.p[ sz( 3 * k0 ) ];
    const double w0 = t.w[ sz( k0 ) ];
    init_box3( c, lo, hi, p0 );
    I64 stack[ 128 ];
    int top = 0;
    stack[ top++ ] = ( I64( 0 ) << 7 ) | t.depth;
    while ( top ) {
        const I64 en = stack[ --top ];
        const I64 m = en >> 7;
        const int h = int( en & 127 );
        const I64 b = t.beg[ sz( m ) ], e = t.end[ sz( m ) ];
        if ( b >= e ) continue;
        if ( ! can_cut3( t, m, c, p0, w0, ev ) ) continue;
        if ( h <= 1 ) {
            for ( I64 k = b; k < e; ++k ) {
                if ( k == k0 ) continue;
                const double dx = t.p[ sz( 3 * k ) ] - p0[ 0 ], dy = t.p[ sz( 3 * k + 1 ) ] - p0[ 1 ], dz = t.p[ sz( 3 * k + 2 ) ] - p0[ 2 ];
                const double off = ( ( w0 - t.w[ sz( k ) ] ) + ( dx * dx + dy * dy + dz * dz ) ) / 2;
                if ( cut3( c, dx, dy, dz, off, int( k ), ev ) == EMPTY ) return false;
            }
            continue;
        }
        const I64 lc = m + 1, rc = m + ( I64( 1 ) << ( h - 1 ) );
        const int hc = h - 1;
        if ( proximity3( t, lc, p0 ) <= proximity3( t, rc, p0 ) ) {
            stack[ top++ ] = ( rc << 7 ) | hc;
            stack[ top++ ] = ( lc << 7 ) | hc;
        } else {
            stack[ top++ ] = ( lc << 7 ) | hc;
            stack[ top++ ] = ( rc << 7 ) | hc;
        }
    }
    return c.nv > 0;
}

/// the same cell cut by every other seed in rank order ( no tree )
bool make_cell3_all( const Tree3 &t, I64 k0, const double *lo, const double *hi, Poly3 &c ) {
    const double *p0 = &t.p[ sz( 3 * k0 ) ];
    const double w0 = t.w[ sz( k0 ) ];
    init_box3( c, lo, hi, p0 );
    for ( I64 k = 0; k < t.n; ++k ) {
        if ( k == k0 ) continue;
        const double dx = t.p[ sz( 3 * k ) ] - p0[ 0 ], dy = t.p[ sz( 3 * k + 1 ) ] - p0[ 1 ], dz = t.p[ sz( 3 * k + 2 ) ] - p0[ 2 ];
        const double off = ( ( w0 - t.w[ sz( k ) ] ) + ( dx * dx + dy * dy + dz * dz ) ) / 2;
        if ( cut3( c, dx, dy, dz, off, int( k ), nullptr ) == EMPTY ) return false;
    }
    return true;
}

/// the mass of the cell and its facets `c_kj = int_face rho / ( 2 |p_j - p_k| )` ( rank `j` )
template<class F>
static double mass_and_facets3( const Tree3 &t, const Density3 &dn, I64 k, const Poly3 &c, bool with_facets, F &&facet ) {
    const double *p0 = &t.p[ sz( 3 * k ) ];
    double m = 0;
    if ( dn.kind == 1 ) {                                // a voxel image: the volumes and face areas of the voxels' pieces
        std::vector<std::pair<int, double>> parts;       // ( the facets' `int rho`, by neighbour rank )
        Density3 unit;
        unit.kind = 0; unit.value = 1;
        for_each_voxel_piece( dn, p0, c, [&]( double rho, const Poly3 &piece ) {
            for_each_face_and_tet( piece, [&]( int f, const std::vector<int> &loop ) {
                const Face3 &fc = piece.f[ sz( f ) ];
                if ( ! with_facets || fc.id < 0 ) return;
                parts.emplace_back( fc.id, rho * face_integral( unit, p0, piece, loop ) );
            }, [&]( const double *, const double *, const double *, const double *, double vol ) { m += rho * vol; } );
        } );
        if ( with_facets ) {
            std::stable_sort( parts.begin(), parts.end(), []( const auto &x, const auto &y ) { return x.first < y.first; } );
            for ( size_t q = 0; q < parts.size(); ) {
                size_t e = q;
                double s = 0;
                while ( e < parts.size() && parts[ e ].first == parts[ q ].first ) s += parts[ e++ ].second;
                const I64 j = parts[ q ].first;
                double d = 0;
                for ( const Face3 &fc : c.f ) if ( fc.id == j ) d = std::sqrt( fc.nx * fc.nx + fc.ny * fc.ny + fc.nz * fc.nz );
                if ( d > 0 ) facet( j, s / ( 2 * d ) );
                q = e;
            }
        }
        return m;
    }
    for_each_face_and_tet( c, [&]( int f, const std::vector<int> &loop ) {
        const Face3 &fc = c.f[ sz( f ) ];
        const double d = std::sqrt( fc.nx * fc.nx + fc.ny * fc.ny + fc.nz * fc.nz );
        if ( dn.kind == 2 && d > 0 ) {                   // the mass: the flux through the face
            const double nrm[ 3 ] = { fc.nx / d, fc.ny / d, fc.nz / d };
            m += face_integral( dn, p0, c, loop, nrm, fc.id < 0 ? BOX_SUB : 1 );
        }
        if ( ! with_facets || fc.id < 0 ) return;
        if ( d > 0 ) facet( I64( fc.id ), face_integral( dn, p0, c, loop ) / ( 2 * d ) );
    }, [&]( const double *, const double *, const double *, const double *, double vol ) {
        if ( dn.kind == 0 ) m += dn.value * vol;
    } );
    return m;
}

// ---- the sweep as Newton reads it ---------------------------------------------------------------------------------

struct Sweep3 {
    static constexpr int DIM = 3;
    Tree3  &t;
    const double *lo, *hi;
    Density3 dn;
    int    nb_diag = 0;
    double t_majorant = 0, t_diag = 0;
    Events ev;
    Edges *edges = nullptr;              ///< ( the limits step is a 2D one: never filled )
    I64    nb_broken = 0;

    Sweep3( Tree3 &t, const double *lo, const double *hi, const Density3 &dn ) : t( t ), lo( lo ), hi( hi ), dn( dn ) {}

    void set_weights( const std::vector<double> &W ) {
        const double t0 = now();
        set_weights3( t, W.data() );
        t_majorant += now() - t0;
    }

    void measures( std::vector<double> &a, std::vector<Facet> *fa ) {
        const double t0 = now();
        const I64 n = t.n;
        a.assign( sz( n ), 0.0 );
        const int nt = nb_threads();
        std::vector<std::vector<Facet>> fth( sz( nt ) );
        std::vector<Events> eth( sz( nt ) );
        std::vector<I64> bth( sz( nt ), 0 );
        parallel_blocks( n, 256, [&]( int th, I64 b, I64 e ) {
            Poly3 c;
            for ( I64 k = b; k < e; ++k ) {
                const I64 i = t.order[ sz( k ) ];
                if ( ! make_cell3( t, k, lo, hi, c, &eth[ sz( th ) ] ) ) { a[ sz( i ) ] = 0; continue; }
                bth[ sz( th ) ] += c.broken;
                a[ sz( i ) ] = mass_and_facets3( t, dn, k, c, fa != nullptr, [&]( I64 j, double cij ) {
                    fth[ sz( th ) ].push_back( Facet{ i, t.order[ sz( j ) ], cij } );
                } );
            }
        } );
        if ( fa ) {
            fa->clear();
            for ( auto &v : fth ) fa->insert( fa->end(), v.begin(), v.end() );
            std::stable_sort( fa->begin(), fa->end(), []( const Facet &x, const Facet &y ) { return x.i < y.i; } );
        }
        for ( auto &e : eth ) ev.add( e );
        for ( I64 v : bth ) nb_broken += v;
        t_diag += now() - t0;
        ++nb_diag;
    }

    /// barycenters ( the seed for an empty cell ) and the transport cost
    void moments( std::vector<double> &bary, double &cost ) {
        const I64 n = t.n;
        bary.assign( sz( n ) * 3, 0.0 );
        const int nt = nb_threads();
        std::vector<double> cth( sz( nt ), 0.0 );
        parallel_blocks( n, 256, [&]( int th, I64 b, I64 e ) {
            Poly3 c;
            for ( I64 k = b; k < e; ++k ) {
                const I64 i = t.order[ sz( k ) ];
                const double *p0 = &t.p[ sz( 3 * k ) ];
                double m = 0, mx[ 3 ] = { 0, 0, 0 }, m2 = 0;
                if ( make_cell3( t, k, lo, hi, c, nullptr ) ) {
                    if ( dn.kind == 1 ) {                // a voxel image: the pieces' moments, each at its voxel's value
                        for_each_voxel_piece( dn, p0, c, [&]( double rho, const Poly3 &piece ) {
                            Density3 cst;
                            cst.kind = 0; cst.value = rho;
                            for_each_face_and_tet( piece, []( int, const std::vector<int> & ) {},
                                                   [&]( const double *A, const double *B, const double *C, const double *D, double vol ) {
                                                       tet_moments( cst, p0, A, B, C, D, vol, m, mx, m2 );
                                                   } );
                        } );
                    } else
                        for_each_face_and_tet( c, []( int, const std::vector<int> & ) {},
                                               [&]( const double *A, const double *B, const double *C, const double *D, double vol ) {
                                                   tet_moments( dn, p0, A, B, C, D, vol, m, mx, m2 );
                                               } );
                }
                cth[ sz( th ) ] += m2;
                for ( int d = 0; d < 3; ++d ) bary[ sz( 3 * i + d ) ] = m > 0 ? p0[ d ] + mx[ d ] / m : p0[ d ];
            }
        } );
        cost = 0;
        for ( double v : cth ) cost += v;
    }
};

/// the similarity start in 3D ( `sdotplan/Solve.h:69-100` ): the cloud mapped into the box by `a y + b`, the weights that
/// make the cells of `pos` those of the mapped cloud
static void similarity3( const double *pos, I64 n, const double *lo, const double *hi, std::vector<double> &w, double margin = 0.1 ) {
    double plo[ 3 ] = { 1e300, 1e300, 1e300 }, phi[ 3 ] = { -1e300, -1e300, -1e300 };
    for ( I64 k = 0; k < n; ++k )
        for ( int d = 0; d < 3; ++d ) { plo[ d ] = std::min( plo[ d ], pos[ 3 * k + d ] ); phi[ d ] = std::max( phi[ d ], pos[ 3 * k + d ] ); }
    double a = 1;
    for ( int d = 0; d < 3; ++d ) a = std::min( a, ( hi[ d ] - lo[ d ] ) * ( 1 - 2 * margin ) / std::max( phi[ d ] - plo[ d ], 1e-300 ) );
    double b[ 3 ];
    for ( int d = 0; d < 3; ++d ) b[ d ] = ( lo[ d ] + hi[ d ] ) / 2 - a * ( plo[ d ] + phi[ d ] ) / 2;
    w.assign( sz( n ), 0.0 );
    for ( I64 k = 0; k < n; ++k ) {
        double pp = 0, qq = 0;
        for ( int d = 0; d < 3; ++d ) {
            const double q = a * pos[ 3 * k + d ] + b[ d ];
            pp += pos[ 3 * k + d ] * pos[ 3 * k + d ];
            qq += q * q;
        }
        w[ sz( k ) ] = pp - qq / a;
    }
}
