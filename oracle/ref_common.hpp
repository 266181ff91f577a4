// Shared pieces of the GenDiL harness (test infrastructure, compiled against
// the reference headers in place by oracle/Makefile into oracle/_ref/; never
// linked into the product library): command-line arguments, the analytic
// coefficient fields of the parity cases, the GenDiL space / rule factories the
// cases use, and the fixture writer and timing loop.
//
// Reference call pattern followed:
//   tests/integration/gendil/test-mass-3d-generic.cpp:83-150
//   tests/integration/gendil/test-advection-3d-generic.cpp:84-170
//   benchmarks/include/benchmark-common.hpp:276-295 (TimeOperator)
#pragma once

#include <gendil/gendil.hpp>

#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

namespace gdref {

using namespace gendil;

struct Args
{
   std::string mode, op, basis, coef, bc;
   int D = 3, P = 3, Q = 4, n = 4, periodic = 0;
   unsigned seed = 12345;
   std::string out;
   int pairs = 1;
};

constexpr double kPi = 3.14159265358979323846;

template < int D >
void advection_field( const std::string & kind, const std::array< Real, D > & X, Real ( &v )[D] )
{
   if ( kind == "c2" )
   {
      if constexpr ( D == 3 )
      {
         v[0] = 1.0 + 0.5 * std::sin( 2.0 * kPi * X[1] );
         v[1] = 0.5 + 0.5 * std::cos( 2.0 * kPi * X[2] );
         v[2] = 0.25 + 0.25 * std::sin( 2.0 * kPi * X[0] );
      }
   }
   else if ( kind == "bubble" )
   {
      for ( int d = 0; d < D; ++d ) v[d] = -X[d] * ( 1.0 - X[d] );
   }
   else if ( kind == "const" )
   {
      Real c = 1.0;
      for ( int d = 0; d < D; ++d ) { v[d] = c; c *= -0.5; }
   }
   else if ( kind == "trig" )
   {
      for ( int d = 0; d < D; ++d ) v[d] = 0.5 + std::sin( 2.0 * kPi * X[( d + 1 ) % D] );
   }
   else if ( kind == "yzx" )
   {
      // tests/integration/gendil/test-advection-p-adaptivity.cpp:62-71
      if constexpr ( D == 3 )
      {
         v[0] = X[1];
         v[1] = X[2];
         v[2] = X[0];
      }
   }
   else if ( kind == "phase" )
   {
      if constexpr ( D == 6 )
      {
         v[0] = X[3];
         v[1] = X[4];
         v[2] = X[5];
         v[3] = 0.1;
         v[4] = 0.0;
         v[5] = 0.0;
      }
   }
   else if ( kind == "ones" )
   {
      for ( int d = 0; d < D; ++d ) v[d] = 1.0;
   }
   else
   {
      std::fprintf( stderr, "unknown advection field %s\n", kind.c_str() );
      std::exit( 2 );
   }
}

template < int D >
Real mass_density( const std::string & kind, const std::array< Real, D > & X )
{
   if ( kind == "one" ) return 1.0;
   if ( kind == "poly" )
   {
      if constexpr ( D >= 2 ) return 1.0 + X[0] * X[1] + X[D - 1] * X[D - 1];
      else return 1.0 + X[0] * X[0];
   }
   std::fprintf( stderr, "unknown mass density %s\n", kind.c_str() );
   std::exit( 2 );
}

template < int D >
Real boundary_value( const std::string & kind, const std::array< Real, D > & X )
{
   if ( kind == "zero" ) return 0.0;
   if ( kind == "poly" )
   {
      if constexpr ( D >= 2 ) return 1.0 + X[0] - 0.5 * X[1];
      else return 1.0 + X[0];
   }
   std::fprintf( stderr, "unknown boundary field %s\n", kind.c_str() );
   std::exit( 2 );
}

template < int D, int P, int Q, bool Lobatto, size_t... Is >
auto make_space( const Args & a, std::index_sequence< Is... > )
{
   std::array< GlobalIndex, D > sizes;
   std::array< Real, D > h;
   Point< D > origin;
   for ( int d = 0; d < D; ++d )
   {
      sizes[d] = a.n;
      h[d] = 1.0 / a.n;
      origin[d] = 0.0;
   }
   CartesianMesh< D > mesh( sizes, h, origin, a.periodic != 0 );
   FiniteElementOrders< ( (void)Is, P )... > orders;
   if constexpr ( Lobatto )
   {
      auto fe = MakeLobattoFiniteElement( orders );
      return MakeFiniteElementSpace( mesh, fe );
   }
   else
   {
      auto fe = MakeLegendreFiniteElement( orders );
      return MakeFiniteElementSpace( mesh, fe );
   }
}

template < int Q, size_t... Is >
auto make_rule( std::index_sequence< Is... > )
{
   IntegrationRuleNumPoints< ( (void)Is, Q )... > nq;
   return MakeIntegrationRule( nq );
}

template < int P, size_t... Is >
auto make_gll( std::index_sequence< Is... > )
{
   return MakeLobattoFiniteElement( FiniteElementOrders< ( (void)Is, P )... >{} );
}

template < typename Op >
double time_pairs( const Op & op, Vector & x, Vector & y, int pairs )
{
   // warm-up pair, then timed pairs (TimeOperator pattern)
   op( x, y );
   op( y, x );
   const auto t0 = std::chrono::steady_clock::now();
   for ( int i = 0; i < pairs; ++i )
   {
      op( x, y );
      op( y, x );
   }
   const auto t1 = std::chrono::steady_clock::now();
   return std::chrono::duration< double >( t1 - t0 ).count() / ( 2.0 * pairs );
}

inline int write_vectors( const std::string & path, const Vector & x, const Vector & y, Integer ndofs )
{
   const Real * xd = x.ReadHostData();
   const Real * yd = y.ReadHostData();
   FILE * f = std::fopen( path.c_str(), "wb" );
   if ( !f ) { std::perror( "fopen" ); return 3; }
   const long long n = ndofs;
   std::fwrite( &n, sizeof( n ), 1, f );
   std::fwrite( xd, sizeof( Real ), ndofs, f );
   std::fwrite( yd, sizeof( Real ), ndofs, f );
   std::fclose( f );
   return 0;
}

// fixture: y = op(x) with x = FillRandom(seed), written as int64 ndofs, x, y.
// bench: warm-up pair and timed pairs of applies, the two vectors swapped each
// apply as in TimeOperator; prints one JSON line.
template < typename Op >
int finish( const Args & a, const Op & op, Integer ndofs, Integer nelem )
{
   Vector x( ndofs );
   Vector y( ndofs );
   FillRandom( x, a.seed );
   y = 0.0;
   if ( a.mode == "fixture" )
   {
      op( x, y );
      return write_vectors( a.out, x, y, ndofs );
   }
   const double t = time_pairs( op, x, y, a.pairs );
   std::printf( "{\"seconds_per_apply\": %.9e, \"dofs\": %lld, \"elements\": %lld, \"dofs_per_s\": %.6e}\n",
                t, (long long)ndofs, (long long)nelem, ndofs / t );
   return 0;
}

// Solve fixture: int64 n, b, x, then iterations, success and the returned
// relative residual.
inline int write_solve( const Args & a, const Vector & b, const Vector & x, Integer ndofs, double iters, bool ok,
                        double relres )
{
   FILE * f = std::fopen( a.out.c_str(), "wb" );
   if ( !f ) { std::perror( "fopen" ); return 3; }
   const long long n = ndofs;
   std::fwrite( &n, sizeof( n ), 1, f );
   std::fwrite( b.ReadHostData(), sizeof( Real ), ndofs, f );
   std::fwrite( x.ReadHostData(), sizeof( Real ), ndofs, f );
   const double tail[3] = { iters, ok ? 1.0 : 0.0, relres };
   std::fwrite( tail, sizeof( double ), 3, f );
   std::fclose( f );
   return 0;
}

// GenDiL's own BiCGSTAB_no_alloc (bicgstab.hpp:44) on K = M + dt A (unit mass
// density, advection field `coef` with boundary field `bc`), b = FillRandom(seed),
// x0 = 0, tol 1e-10, at most 500 iterations.  Writes n, b, x, iterations,
// success and the returned relative residual.
template < int D, typename Space, typename Rule >
int solve_bicgstab( const Args & a, const Space & fes, const Rule & ir, Integer ndofs )
{
   const std::string coef = a.coef, bc = a.bc;
   auto sigma = []( const std::array< Real, D > & ) -> Real { return 1.0; };
   auto adv = [coef]( const std::array< Real, D > & X, Real ( &v )[D] ) { advection_field< D >( coef, X, v ); };
   auto g = [bc]( const std::array< Real, D > & X ) -> Real { return boundary_value< D >( bc, X ); };
   auto M = MakeMassFiniteElementOperator< SerialKernelConfiguration >( fes, ir, sigma );
   auto A = MakeAdvectionOperator< SerialKernelConfiguration >( fes, ir, adv, g );
   const Real dt = 0.05;
   Vector tmp( ndofs );
   auto K = [&]( const Vector & x, Vector & y )
   {
      M( x, y );
      A( x, tmp );
      Axpy( dt, tmp, y );
   };
   auto dot = []( const Vector & u, const Vector & v ) { return Dot( u, v ); };
   Vector b( ndofs ), x( ndofs ), r( ndofs ), r_hat( ndofs ), p( ndofs ), v( ndofs ), s( ndofs ), t( ndofs );
   FillRandom( b, a.seed );
   x = 0.0;
   auto [ok, iters, relres] = BiCGSTAB_no_alloc( K, b, dot, 500, 1e-10, x, r, r_hat, p, v, s, t );
   return write_solve( a, b, x, ndofs, (double)iters, ok, (double)relres );
}


// GenDiL's own GMRES_no_alloc (gmres.hpp:66-288) on K = M + dt A as in
// solve_bicgstab, restart 30, b = FillRandom(seed), x0 = 0, tol 1e-10, at
// most 500 iterations.  Same output as solve_bicgstab.
template < int D, typename Space, typename Rule >
int solve_gmres( const Args & a, const Space & fes, const Rule & ir, Integer ndofs )
{
   const std::string coef = a.coef, bc = a.bc;
   auto sigma = []( const std::array< Real, D > & ) -> Real { return 1.0; };
   auto adv = [coef]( const std::array< Real, D > & X, Real ( &v )[D] ) { advection_field< D >( coef, X, v ); };
   auto g = [bc]( const std::array< Real, D > & X ) -> Real { return boundary_value< D >( bc, X ); };
   auto M = MakeMassFiniteElementOperator< SerialKernelConfiguration >( fes, ir, sigma );
   auto A = MakeAdvectionOperator< SerialKernelConfiguration >( fes, ir, adv, g );
   const Real dt = 0.05;
   Vector tmp( ndofs );
   auto K = [&]( const Vector & x, Vector & y )
   {
      M( x, y );
      A( x, tmp );
      Axpy( dt, tmp, y );
   };
   auto dot = []( const Vector & u, const Vector & v ) { return Dot( u, v ); };
   const Integer restart = 30;
   std::vector< Vector > V;
   for ( Integer i = 0; i <= restart; ++i ) V.emplace_back( ndofs );
   std::vector< Real > H( ( restart + 1 ) * restart ), cs( restart ), sn( restart ), e1( restart + 1 ), yv( restart );
   Vector b( ndofs ), x( ndofs ), w( ndofs );
   FillRandom( b, a.seed );
   x = 0.0;
   auto [ok, iters, relres] = GMRES_no_alloc( K, b, dot, 500, restart, 1e-10, x, V, w, H, cs, sn, e1, yv );
   return write_solve( a, b, x, ndofs, (double)iters, ok, (double)relres );
}

// GenDiL's own ConjugateGradient (conjugategradient.hpp:118-208, the variant
// with an injected dot and scratch vectors) on the mass matrix M(sigma = coef),
// b = FillRandom(seed), x0 = 0, tol 1e-10, at most 500 iterations.
template < int D, typename Space, typename Rule >
int solve_cg( const Args & a, const Space & fes, const Rule & ir, Integer ndofs )
{
   const std::string coef = a.coef;
   auto sigma = [coef]( const std::array< Real, D > & X ) -> Real { return mass_density< D >( coef, X ); };
   auto M = MakeMassFiniteElementOperator< SerialKernelConfiguration >( fes, ir, sigma );
   auto dot = []( const Vector & u, const Vector & v ) { return Dot( u, v ); };
   Vector b( ndofs ), x( ndofs ), tmp( ndofs ), z( ndofs ), r( ndofs ), p( ndofs );
   FillRandom( b, a.seed );
   x = 0.0;
   auto [ok, iters, relres] = ConjugateGradient( M, b, dot, 500, 1e-10, x, tmp, z, r, p );
   return write_solve( a, b, x, ndofs, (double)iters, ok, (double)relres );
}

template < int D, int P, int Q, bool Lobatto >
int run( const Args & a )
{
   auto fes = make_space< D, P, Q, Lobatto >( a, std::make_index_sequence< D >{} );
   auto ir = make_rule< Q >( std::make_index_sequence< D >{} );
   const Integer ndofs = fes.GetNumberOfFiniteElementDofs();
   const Integer nelem = fes.GetNumberOfFiniteElements();
   const std::string coef = a.coef;
   const std::string bc = a.bc;
   if ( a.mode == "bicgstab" ) return solve_bicgstab< D >( a, fes, ir, ndofs );
   if ( a.mode == "gmres" ) return solve_gmres< D >( a, fes, ir, ndofs );
   if ( a.mode == "cg" ) return solve_cg< D >( a, fes, ir, ndofs );
   if ( a.op == "mass" )
   {
      auto sigma = [coef]( const std::array< Real, D > & X ) -> Real { return mass_density< D >( coef, X ); };
      auto op = MakeMassFiniteElementOperator< SerialKernelConfiguration >( fes, ir, sigma );
      return finish( a, op, ndofs, nelem );
   }
   if ( a.op == "adv" )
   {
      auto adv = [coef]( const std::array< Real, D > & X, Real ( &v )[D] ) { advection_field< D >( coef, X, v ); };
      if ( bc == "none" )
      {
         auto op = MakeAdvectionOperator< SerialKernelConfiguration >( fes, ir, adv );
         return finish( a, op, ndofs, nelem );
      }
      auto g = [bc]( const std::array< Real, D > & X ) -> Real { return boundary_value< D >( bc, X ); };
      auto op = MakeAdvectionOperator< SerialKernelConfiguration >( fes, ir, adv, g );
      return finish( a, op, ndofs, nelem );
   }
   std::fprintf( stderr, "unknown op %s\n", a.op.c_str() );
   return 2;
}

using Runner = int ( * )( const Args & );

struct CaseEntry { int D, P, Q; char basis; Runner fn; };

#define GD_CASE( D, P, Q ) { D, P, Q, 'L', &::gdref::run< D, P, Q, true > }
#define GD_CASE_G( D, P, Q ) { D, P, Q, 'G', &::gdref::run< D, P, Q, false > }

// Entry points of the other translation units of the harness; each returns 2
// when the requested case is not instantiated.
int run_padapt_case( const Args & a );
// mass / adv / bicgstab cases of 5D meshes (ref_highdim.cpp) and of the (n, q)
// pairs of GenDiL's range benchmarks in 3D, 4D and 6D (ref_rangebench.cpp)
int run_highdim_case( const Args & a );
int run_rangebench_case( const Args & a );
// SIPDG diffusion (ref_diffusion.cpp)
int run_diffusion_case( const Args & a );

} // namespace gdref
