// The oracle's multigrid for the Newton systems: the linear solver the reference's CPU solve picks by default when it is
// built without Eigen ( `sdotplan/Linear.cpp:208-233`: AUTO -> MG, `nu = 1` in 3D ), restated for the CPU baseline and
// for its own tests. TEST INFRASTRUCTURE: only the tests, the smoke check and the bench's CPU leg use it.
//
// What it follows ( `sdotplan/Multigrid.h:1-729` ), in code of its own:
// * aggregation by index arithmetic: the packet of a seed is its tree rank >> 3 at the first level, its index >> 3 below;
// * smoothed prolongation `P = ( I - 0.7 D^-1 A ) P0`, each row truncated to the entries of at least 0.2 of its largest and
//   renormalized to a unit sum ( `Multigrid.h:331-368` ), Galerkin `A_c = P^t A P` ( `Multigrid.h:399-459` );
// * Chebyshev smoothing of degree `nu` ( 3 in 2D, 1 in 3D ) in `M^-1 A`, `M` the SPAI-0 diagonal `A_ii / sum_j A_ij^2`,
//   on `[ lmax / 10, lmax ]` with `lmax` the Gershgorin bound ( `Multigrid.h:238-302` ); the recurrence is the textbook
//   Chebyshev acceleration ( Saad, Iterative Methods for Sparse Linear Systems, 2nd ed., Algorithm 12.1 );
// * the coarsest level ( <= 1000 unknowns ) smoothed 120 times from zero ( no factorization: the build without Eigen );
// * a V-cycle as the preconditioner of CG in the zero-mean gauge; the hierarchy serves 4 solves ( only the finest level's
//   matrix and coefficients are refreshed in between ); the last 2 solutions start the next solve by a Galerkin projection
//   ( `Multigrid.h:614-680` ); the result is returned in the gauge `d[ 0 ] = 0`.
//
// Loops run in contiguous slices over a small persistent pool of `nb_threads()` workers; sums are combined in slice order, so
// a result depends on the thread count only through those sums' rounding.
#pragma once

// ( included inside the oracle's namespace, after `Laplacian` and `LinStats`; `<condition_variable>`, `<mutex>` and
// `<memory>` are included at the top of its translation unit )


/// a persistent pool: `run( body )` calls `body( t )` for `t` in `[ 0, size )`, the caller taking `t = 0`
class MgPool {
public:
    explicit MgPool( int nt ) : nt_( std::max( nt, 1 ) ) {
        for ( int t = 1; t < nt_; ++t )
            th_.emplace_back( [this, t] { loop( t ); } );
    }
    ~MgPool() {
        { std::lock_guard<std::mutex> g( mu_ ); quit_ = true; ++gen_; }
        cv_.notify_all();
        for ( auto &x : th_ ) x.join();
    }
    int size() const { return nt_; }
    void run( const std::function<void( int )> &body ) {
        if ( nt_ == 1 ) { body( 0 ); return; }
        {
            std::lock_guard<std::mutex> g( mu_ );
            job_ = &body;
            left_ = nt_ - 1;
            ++gen_;
        }
        cv_.notify_all();
        body( 0 );
        std::unique_lock<std::mutex> g( mu_ );
        done_.wait( g, [this] { return left_ == 0; } );
        job_ = nullptr;
    }
private:
    void loop( int t ) {
        unsigned seen = 0;
        for ( ;; ) {
            const std::function<void( int )> *job;
            {
                std::unique_lock<std::mutex> g( mu_ );
                cv_.wait( g, [&] { return gen_ != seen; } );
                seen = gen_;
                if ( quit_ ) return;
                job = job_;
            }
            ( *job )( t );
            std::lock_guard<std::mutex> g( mu_ );
            if ( --left_ == 0 ) done_.notify_one();
        }
    }
    int nt_;
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable cv_, done_;
    const std::function<void( int )> *job_ = nullptr;
    unsigned gen_ = 0;
    int left_ = 0;
    bool quit_ = false;
};

/// the multigrid-preconditioned CG of a Newton solve ( one object per solve: it keeps the hierarchy and the subspace )
class MgSolver {
public:
    int    nu = 3, recycle = 2, rebuild = 4, bottom_sweeps = 120;
    I64    stop = 1000;
    double omega = 0.7, truncate = 0.2, band = 10;

    explicit MgSolver( int threads ) : pool_( threads ) {}

    /// the tree rank of each user index: the first level's packets are `rank >> 3`
    void set_ranks( const std::vector<I64> &rank_of ) { rank_ = rank_of; levels_.clear(); basis_.clear(); }

    /// `L d = b` to `|r| <= tol |b|` ( zero-mean gauge inside ), `d[ 0 ] = 0` on return; false if not converged to 1
    bool solve( const Laplacian &L, const std::vector<double> &b, std::vector<double> &d, double tol, int maxit, LinStats &st ) {
        const double t0 = now();
        if ( ! levels_.empty() && levels_[ 0 ].n != L.n ) basis_.clear();
        if ( levels_.empty() || levels_[ 0 ].n != L.n || levels_.size() == 1 || since_ >= rebuild ) { build( L ); since_ = 0; }
        else { point_at( levels_[ 0 ], L ); coefficients( levels_[ 0 ] ); }
        ++since_;
        const bool ok = pcg( L, b, d, tol, maxit, st );
        st.t += now() - t0;
        return ok;
    }

    int nb_levels() const { return int( levels_.size() ); }

private:
    struct Csr {
        std::vector<I64> row, col;
        std::vector<double> val;
    };
    /// `y_i = dia_i x_i - sum val x_col`, the Laplacian's own convention
    struct Level {
        I64 n = 0;
        const I64 *row = nullptr, *col = nullptr;
        const double *val = nullptr, *dia = nullptr;
        Csr own;
        std::vector<double> own_dia;
        std::vector<double> spai, x, b, r, y, z;
        double lmax = 2;
    };
    MgPool pool_;
    std::vector<I64> rank_;
    std::vector<Level> levels_;
    std::vector<Csr> prol_, restr_;
    std::vector<std::vector<double>> basis_;
    int since_ = 0;

    int slices( I64 n ) const { return n < 4096 ? 1 : pool_.size(); }

    /// `body( i, t )` for `i < n`, contiguous slices over the pool, `t` the slice
    template<class F> void each_t( I64 n, F &&body ) {
        const int T = slices( n );
        if ( T == 1 ) { for ( I64 i = 0; i < n; ++i ) body( i, 0 ); return; }
        pool_.run( [&]( int t ) {
            const I64 e = n * ( t + 1 ) / T;
            for ( I64 i = n * t / T; i < e; ++i ) body( i, t );
        } );
    }
    template<class F> void each( I64 n, F &&body ) { each_t( n, [&]( I64 i, int ) { body( i ); } ); }
    template<class F> double sum( I64 n, F &&term ) {
        const int T = slices( n );
        std::vector<double> part( static_cast<size_t>( T ), 0.0 );
        auto slice = [&]( int t ) {
            const I64 e = n * ( t + 1 ) / T;
            double s = 0;
            for ( I64 i = n * t / T; i < e; ++i ) s += term( i );
            part[ size_t( t ) ] = s;
        };
        if ( T == 1 ) slice( 0 ); else pool_.run( slice );
        double s = 0;
        for ( double p : part ) s += p;
        return s;
    }

    static void point_at( Level &v, const Laplacian &L ) {
        v.n = L.n;
        v.row = L.row.data(); v.col = L.col.data(); v.val = L.c.data(); v.dia = L.dia.data();
    }

    /// SPAI-0 diagonal and the Gershgorin bound of `M^-1 A`
    void coefficients( Level &v ) {
        v.spai.resize( size_t( v.n ) );
        each( v.n, [&]( I64 i ) {
            double q = v.dia[ i ] * v.dia[ i ];
            for ( I64 e = v.row[ i ]; e < v.row[ i + 1 ]; ++e ) q += v.val[ e ] * v.val[ e ];
            v.spai[ size_t( i ) ] = q > 0 ? v.dia[ i ] / q : 0.0;
        } );
        double mx = 0;
        for ( I64 i = 0; i < v.n; ++i ) {
            double s = v.dia[ i ];
            for ( I64 e = v.row[ i ]; e < v.row[ i + 1 ]; ++e ) s += std::fabs( v.val[ e ] );
            mx = std::max( mx, v.spai[ size_t( i ) ] * s );
        }
        v.lmax = mx > 0 ? mx : 2.0;
    }

    /// `P^t` by a counting transpose ( rows of `P^t` in increasing fine index )
    static void transpose( const Csr &P, I64 nf, I64 nc, Csr &T ) {
        T.row.assign( size_t( nc + 1 ), 0 );
        for ( I64 e = 0; e < P.row[ size_t( nf ) ]; ++e ) ++T.row[ size_t( P.col[ size_t( e ) ] + 1 ) ];
        for ( I64 a = 0; a < nc; ++a ) T.row[ size_t( a + 1 ) ] += T.row[ size_t( a ) ];
        T.col.resize( size_t( T.row[ size_t( nc ) ] ) );
        T.val.resize( T.col.size() );
        std::vector<I64> at( T.row.begin(), T.row.end() - 1 );
        for ( I64 i = 0; i < nf; ++i )
            for ( I64 e = P.row[ size_t( i ) ]; e < P.row[ size_t( i + 1 ) ]; ++e ) {
                const I64 p = at[ size_t( P.col[ size_t( e ) ] ) ]++;
                T.col[ size_t( p ) ] = i;
                T.val[ size_t( p ) ] = P.val[ size_t( e ) ];
            }
    }

    /// a dense accumulator over the columns of one row: `add( col, value )`, `flush` appends the touched entries in
    /// first-touch order and clears them
    struct Acc {
        std::vector<double> val;
        std::vector<char> used;
        std::vector<I64> touched;
        void reset( I64 nc ) { val.assign( size_t( nc ), 0.0 ); used.assign( size_t( nc ), 0 ); touched.clear(); }
        void add( I64 c, double v ) {
            if ( ! used[ size_t( c ) ] ) { used[ size_t( c ) ] = 1; touched.push_back( c ); }
            val[ size_t( c ) ] += v;
        }
    };
    std::vector<Acc> acc_;

    /// the rows of a sparse matrix over `nc` columns, each built by `make_row( i, acc, cols, vals )` in the arena of its
    /// slice, then concatenated
    template<class F> void rows( I64 nr, I64 nc, Csr &out, F &&make_row ) {
        const int T = slices( nr );
        acc_.resize( size_t( T ) );
        for ( int t = 0; t < T; ++t ) acc_[ size_t( t ) ].reset( nc );
        std::vector<std::vector<I64>> tc( static_cast<size_t>( T ) );
        std::vector<std::vector<double>> tv( static_cast<size_t>( T ) );
        std::vector<I64> at( static_cast<size_t>( nr ) ), len( static_cast<size_t>( nr ) );
        std::vector<int> thr( static_cast<size_t>( nr ) );
        each_t( nr, [&]( I64 i, int t ) {
            const size_t b = tc[ size_t( t ) ].size();
            make_row( i, acc_[ size_t( t ) ], tc[ size_t( t ) ], tv[ size_t( t ) ] );
            at[ size_t( i ) ] = I64( b );
            len[ size_t( i ) ] = I64( tc[ size_t( t ) ].size() - b );
            thr[ size_t( i ) ] = t;
        } );
        out.row.assign( size_t( nr + 1 ), 0 );
        for ( I64 i = 0; i < nr; ++i ) out.row[ size_t( i + 1 ) ] = out.row[ size_t( i ) ] + len[ size_t( i ) ];
        out.col.resize( size_t( out.row[ size_t( nr ) ] ) );
        out.val.resize( out.col.size() );
        each( nr, [&]( I64 i ) {
            const int t = thr[ size_t( i ) ];
            for ( I64 k = 0; k < len[ size_t( i ) ]; ++k ) {
                out.col[ size_t( out.row[ size_t( i ) ] + k ) ] = tc[ size_t( t ) ][ size_t( at[ size_t( i ) ] + k ) ];
                out.val[ size_t( out.row[ size_t( i ) ] + k ) ] = tv[ size_t( t ) ][ size_t( at[ size_t( i ) ] + k ) ];
            }
        } );
    }

    void build( const Laplacian &L ) {
        levels_.clear(); prol_.clear(); restr_.clear();
        levels_.emplace_back();
        point_at( levels_[ 0 ], L );
        for ( int l = 0; levels_[ size_t( l ) ].n > stop && l < 24; ++l ) coarsen( l );
        for ( Level &v : levels_ ) {
            for ( auto *p : { &v.x, &v.b, &v.r, &v.y, &v.z } ) p->assign( size_t( v.n ), 0.0 );
            coefficients( v );
        }
    }

    void coarsen( int l ) {
        const Level &f = levels_[ size_t( l ) ];
        const I64 nf = f.n, nc = ( nf + 7 ) / 8;
        std::vector<I64> agg( static_cast<size_t>( nf ) );
        for ( I64 i = 0; i < nf; ++i ) agg[ size_t( i ) ] = ( l == 0 && ! rank_.empty() ? rank_[ size_t( i ) ] : i ) >> 3;
        // P: row i = ( 1 - w ) e_agg( i ) + ( w / sum_j c_ij ) sum_j c_ij e_agg( j ), truncated, renormalized
        Csr P;
        rows( nf, nc, P, [&]( I64 i, Acc &a, std::vector<I64> &cols, std::vector<double> &vals ) {
            double cs = 0;
            for ( I64 e = f.row[ i ]; e < f.row[ i + 1 ]; ++e ) cs += f.val[ e ];
            a.add( agg[ size_t( i ) ], 1 - omega );
            const double k = cs > 0 ? omega / cs : 0.0;
            for ( I64 e = f.row[ i ]; e < f.row[ i + 1 ]; ++e ) a.add( agg[ size_t( f.col[ e ] ) ], k * f.val[ e ] );
            double big = 0;
            for ( I64 c : a.touched ) big = std::max( big, std::fabs( a.val[ size_t( c ) ] ) );
            const size_t b = cols.size();
            double s = 0;
            for ( I64 c : a.touched ) {
                const double v = a.val[ size_t( c ) ];
                a.val[ size_t( c ) ] = 0; a.used[ size_t( c ) ] = 0;
                if ( v != 0 && std::fabs( v ) >= truncate * big ) { cols.push_back( c ); vals.push_back( v ); s += v; }
            }
            a.touched.clear();
            if ( s > 0 ) for ( size_t q = b; q < vals.size(); ++q ) vals[ q ] /= s;
        } );
        Csr R;
        transpose( P, nf, nc, R );
        // A P, one fine row at a time: dia_i P_i. - sum_j c_ij P_j.
        Csr AP;
        rows( nf, nc, AP, [&]( I64 i, Acc &a, std::vector<I64> &cols, std::vector<double> &vals ) {
            for ( I64 e = P.row[ size_t( i ) ]; e < P.row[ size_t( i + 1 ) ]; ++e ) a.add( P.col[ size_t( e ) ], f.dia[ i ] * P.val[ size_t( e ) ] );
            for ( I64 e = f.row[ i ]; e < f.row[ i + 1 ]; ++e ) {
                const I64 j = f.col[ e ];
                for ( I64 g = P.row[ size_t( j ) ]; g < P.row[ size_t( j + 1 ) ]; ++g ) a.add( P.col[ size_t( g ) ], -f.val[ e ] * P.val[ size_t( g ) ] );
            }
            for ( I64 c : a.touched ) { cols.push_back( c ); vals.push_back( a.val[ size_t( c ) ] ); a.val[ size_t( c ) ] = 0; a.used[ size_t( c ) ] = 0; }
            a.touched.clear();
        } );
        // A_c = P^t ( A P ), one coarse row at a time; the diagonal apart, the off-diagonal entries negated
        Level c;
        c.n = nc;
        c.own_dia.assign( size_t( nc ), 1.0 );
        rows( nc, nc, c.own, [&]( I64 r, Acc &a, std::vector<I64> &cols, std::vector<double> &vals ) {
            for ( I64 k = R.row[ size_t( r ) ]; k < R.row[ size_t( r + 1 ) ]; ++k ) {
                const I64 i = R.col[ size_t( k ) ];
                for ( I64 g = AP.row[ size_t( i ) ]; g < AP.row[ size_t( i + 1 ) ]; ++g ) a.add( AP.col[ size_t( g ) ], R.val[ size_t( k ) ] * AP.val[ size_t( g ) ] );
            }
            for ( I64 q : a.touched ) {
                const double v = a.val[ size_t( q ) ];
                a.val[ size_t( q ) ] = 0; a.used[ size_t( q ) ] = 0;
                if ( q == r ) { if ( v > 0 ) c.own_dia[ size_t( r ) ] = v; continue; }
                cols.push_back( q ); vals.push_back( -v );
            }
            a.touched.clear();
        } );
        prol_.push_back( std::move( P ) );
        restr_.push_back( std::move( R ) );
        levels_.push_back( std::move( c ) );
        Level &v = levels_.back();
        v.row = v.own.row.data(); v.col = v.own.col.data(); v.val = v.own.val.data(); v.dia = v.own_dia.data();
    }

    /// `r = b - A x`
    void residual( Level &v ) {
        each( v.n, [&]( I64 i ) {
            double s = v.dia[ i ] * v.x[ size_t( i ) ];
            for ( I64 e = v.row[ i ]; e < v.row[ i + 1 ]; ++e ) s -= v.val[ e ] * v.x[ size_t( v.col[ e ] ) ];
            v.r[ size_t( i ) ] = v.b[ size_t( i ) ] - s;
        } );
    }

    /// `deg` steps of Chebyshev acceleration of `x <- x + M ( b - A x )` on `[ lmax / band, lmax ]`; `from_zero`: `x = 0` first
    void smooth( Level &v, int deg, bool from_zero ) {
        if ( from_zero ) { std::fill( v.x.begin(), v.x.end(), 0.0 ); v.r = v.b; }
        else residual( v );
        if ( deg <= 0 ) return;
        const double hi = v.lmax, lo = hi / band;
        const double theta = 0.5 * ( hi + lo ), delta = 0.5 * ( hi - lo ), sigma = theta / delta;
        double rho = 1 / sigma;
        each( v.n, [&]( I64 i ) { v.y[ size_t( i ) ] = v.spai[ size_t( i ) ] * v.r[ size_t( i ) ] / theta; } );
        for ( int k = 0; ; ++k ) {
            each( v.n, [&]( I64 i ) { v.x[ size_t( i ) ] += v.y[ size_t( i ) ]; } );
            if ( k + 1 >= deg ) break;
            const double rho2 = 1 / ( 2 * sigma - rho ), a = rho2 * rho, b = 2 * rho2 / delta;
            // r -= A y, and the next direction into z ( y is still read by the neighbours' rows )
            each( v.n, [&]( I64 i ) {
                double s = v.dia[ i ] * v.y[ size_t( i ) ];
                for ( I64 e = v.row[ i ]; e < v.row[ i + 1 ]; ++e ) s -= v.val[ e ] * v.y[ size_t( v.col[ e ] ) ];
                const double ri = v.r[ size_t( i ) ] - s;
                v.r[ size_t( i ) ] = ri;
                v.z[ size_t( i ) ] = a * v.y[ size_t( i ) ] + b * v.spai[ size_t( i ) ] * ri;
            } );
            v.y.swap( v.z );
            rho = rho2;
        }
    }

    void vcycle( size_t l ) {
        Level &f = levels_[ l ];
        if ( l + 1 == levels_.size() ) { smooth( f, bottom_sweeps, true ); return; }
        Level &c = levels_[ l + 1 ];
        smooth( f, nu, true );
        residual( f );
        const Csr &R = restr_[ l ];
        each( c.n, [&]( I64 a ) {
            double s = 0;
            for ( I64 k = R.row[ size_t( a ) ]; k < R.row[ size_t( a + 1 ) ]; ++k ) s += R.val[ size_t( k ) ] * f.r[ size_t( R.col[ size_t( k ) ] ) ];
            c.b[ size_t( a ) ] = s;
        } );
        vcycle( l + 1 );
        const Csr &P = prol_[ l ];
        each( f.n, [&]( I64 i ) {
            double s = 0;
            for ( I64 e = P.row[ size_t( i ) ]; e < P.row[ size_t( i + 1 ) ]; ++e ) s += P.val[ size_t( e ) ] * c.x[ size_t( P.col[ size_t( e ) ] ) ];
            f.x[ size_t( i ) ] += s;
        } );
        smooth( f, nu, false );
    }

    void center( std::vector<double> &v ) {
        const I64 n = I64( v.size() );
        const double mu = sum( n, [&]( I64 i ) { return v[ size_t( i ) ]; } ) / double( n );
        each( n, [&]( I64 i ) { v[ size_t( i ) ] -= mu; } );
    }

    void product( const Laplacian &L, const std::vector<double> &x, std::vector<double> &y ) {
        each( L.n, [&]( I64 i ) {
            double s = L.dia[ size_t( i ) ] * x[ size_t( i ) ];
            for ( I64 e = L.row[ size_t( i ) ]; e < L.row[ size_t( i + 1 ) ]; ++e ) s -= L.c[ size_t( e ) ] * x[ size_t( L.col[ size_t( e ) ] ) ];
            y[ size_t( i ) ] = s;
        } );
    }
    double dot( const std::vector<double> &u, const std::vector<double> &v ) {
        return sum( I64( u.size() ), [&]( I64 i ) { return u[ size_t( i ) ] * v[ size_t( i ) ]; } );
    }

    /// `G y = f` by Cholesky with a ridge of `1e-12` of the mean diagonal; false if `G` is not positive definite
    static bool small_spd( std::vector<double> &G, std::vector<double> &f, int k ) {
        double tr = 0;
        for ( int j = 0; j < k; ++j ) tr += G[ size_t( j * k + j ) ];
        if ( ! ( tr > 0 ) ) return false;
        for ( int j = 0; j < k; ++j ) G[ size_t( j * k + j ) ] += 1e-12 * tr / k;
        for ( int j = 0; j < k; ++j ) {
            for ( int q = 0; q < j; ++q ) G[ size_t( j * k + j ) ] -= G[ size_t( j * k + q ) ] * G[ size_t( j * k + q ) ];
            if ( ! ( G[ size_t( j * k + j ) ] > 0 ) ) return false;
            G[ size_t( j * k + j ) ] = std::sqrt( G[ size_t( j * k + j ) ] );
            for ( int i = j + 1; i < k; ++i ) {
                for ( int q = 0; q < j; ++q ) G[ size_t( i * k + j ) ] -= G[ size_t( i * k + q ) ] * G[ size_t( j * k + q ) ];
                G[ size_t( i * k + j ) ] /= G[ size_t( j * k + j ) ];
            }
        }
        for ( int i = 0; i < k; ++i ) { for ( int q = 0; q < i; ++q ) f[ size_t( i ) ] -= G[ size_t( i * k + q ) ] * f[ size_t( q ) ]; f[ size_t( i ) ] /= G[ size_t( i * k + i ) ]; }
        for ( int i = k - 1; i >= 0; --i ) { for ( int q = i + 1; q < k; ++q ) f[ size_t( i ) ] -= G[ size_t( q * k + i ) ] * f[ size_t( q ) ]; f[ size_t( i ) ] /= G[ size_t( i * k + i ) ]; }
        return true;
    }

    bool pcg( const Laplacian &L, const std::vector<double> &b, std::vector<double> &d, double tol, int maxit, LinStats &st ) {
        const I64 n = L.n;
        std::vector<double> r( b.begin(), b.begin() + n ), z( static_cast<size_t>( n ) ), p( static_cast<size_t>( n ) ), q( static_cast<size_t>( n ) );
        d.assign( size_t( n ), 0.0 );
        center( r );
        const double bb = dot( r, r );
        if ( ! ( bb > 0 ) ) return true;
        // the start: the energy-best combination of the kept solutions, and its residual
        if ( recycle > 0 && ! basis_.empty() ) {
            const int k = int( basis_.size() );
            std::vector<std::vector<double>> AU( static_cast<size_t>( k ), std::vector<double>( static_cast<size_t>( n ) ) );
            for ( int j = 0; j < k; ++j ) product( L, basis_[ size_t( j ) ], AU[ size_t( j ) ] );
            std::vector<double> G( static_cast<size_t>( k * k ) ), f( static_cast<size_t>( k ) );
            for ( int j = 0; j < k; ++j ) {
                f[ size_t( j ) ] = dot( basis_[ size_t( j ) ], r );
                for ( int i = 0; i <= j; ++i ) G[ size_t( j * k + i ) ] = G[ size_t( i * k + j ) ] = dot( basis_[ size_t( j ) ], AU[ size_t( i ) ] );
            }
            if ( small_spd( G, f, k ) ) {
                each( n, [&]( I64 i ) {
                    double sx = 0, sr = 0;
                    for ( int j = 0; j < k; ++j ) { sx += f[ size_t( j ) ] * basis_[ size_t( j ) ][ size_t( i ) ]; sr += f[ size_t( j ) ] * AU[ size_t( j ) ][ size_t( i ) ]; }
                    d[ size_t( i ) ] = sx;
                    r[ size_t( i ) ] -= sr;
                } );
                center( d );
                center( r );
            }
        }
        auto precondition = [&]() {
            levels_[ 0 ].b = r;
            vcycle( 0 );
            z = levels_[ 0 ].x;
            center( z );
        };
        precondition();
        p = z;
        double rz = dot( r, z ), rr = dot( r, r );
        const double goal = tol * tol * bb;
        int it = 0;
        while ( it < maxit && rr > goal ) {
            product( L, p, q );
            const double pq = dot( p, q );
            if ( ! ( pq > 0 ) ) break;
            const double alpha = rz / pq;
            rr = sum( n, [&]( I64 i ) {
                d[ size_t( i ) ] += alpha * p[ size_t( i ) ];
                r[ size_t( i ) ] -= alpha * q[ size_t( i ) ];
                return r[ size_t( i ) ] * r[ size_t( i ) ];
            } );
            ++it;
            if ( rr <= goal ) break;
            precondition();
            const double rz2 = dot( r, z ), beta = rz != 0 ? rz2 / rz : 0.0;
            rz = rz2;
            each( n, [&]( I64 i ) { p[ size_t( i ) ] = z[ size_t( i ) ] + beta * p[ size_t( i ) ]; } );
        }
        st.nb_iter += it;
        if ( recycle > 0 ) {
            if ( int( basis_.size() ) >= recycle ) basis_.erase( basis_.begin() );
            basis_.push_back( d );
        }
        const double d0 = d[ 0 ];
        each( n, [&]( I64 i ) { d[ size_t( i ) ] -= d0; } );
        const double err = std::sqrt( rr / bb );
        st.worst = std::max( st.worst, err );
        return err < 1;
    }
};
