// SIPDG diffusion cases of the GenDiL harness (test infrastructure; see
// ref_harness.cpp): GenDiL's DiffusionOperator (diffusionoperator.hpp:44-331,
// factory :519-537) on CartesianMesh<D>, called as its own tests and examples
// call it:
//   tests/integration/gendil/test-diffusion-3d-generic.cpp:78-167 (GLL,
//     q = p + 3, n = 2 periodic, sigma = -1, kappa = (p + 1)^2)
//   tests/integration/gendil/test-diffusion-symmetry-1d-minimal.cpp:10-40
//     (1D, p = 1, q = 4, sigma = 1, kappa = 0)
//   examples/poisson/poisson-convergence-{1d,2d,6d}.cpp (Legendre elements,
//     q = p + 2, sigma = -1, kappa = (p + 1)^2, non-periodic meshes)
//
// Command line: op "diff", coef = diffusivity kind, bc = "s<sigma>_k<kappa>".
#include "ref_common.hpp"

namespace gdref {

namespace {

// Diffusivity fields.  "trig" is the field of test-diffusion-3d-generic.cpp:
// 121-133 without its last term (0.05 cos(2 pi (3x - y + 2z))): with it the
// field is a sum of 12 separable terms, more than a separable field
// descriptor carries (8).
template < int D >
Real diffusivity( const std::string & kind, const std::array< Real, D > & X )
{
   if ( kind == "one" ) return 1.0;
   if ( kind == "poly" )
   {
      return 1.0 + X[0] * X[D - 1] + 0.5 * X[0] * X[0];
   }
   if ( kind == "trig" )
   {
      const Real two_pi = 2.0 * kPi;
      if constexpr ( D == 3 )
      {
         return 2.0 + 0.25 * std::sin( two_pi * X[0] ) + 0.17 * std::cos( two_pi * ( 2.0 * X[1] + 0.13 ) ) +
                0.11 * std::sin( two_pi * ( 3.0 * X[2] + 0.37 ) ) +
                0.07 * std::sin( two_pi * ( X[0] + 2.0 * X[1] + 3.0 * X[2] ) );
      }
      else
      {
         return 2.0 + 0.25 * std::sin( two_pi * X[0] ) + 0.17 * std::cos( two_pi * ( 2.0 * X[D - 1] + 0.13 ) );
      }
   }
   std::fprintf( stderr, "unknown diffusivity %s\n", kind.c_str() );
   std::exit( 2 );
}

template < int D, int P, int Q, bool Lobatto >
int run_diff( const Args & a )
{
   auto fes = make_space< D, P, Q, Lobatto >( a, std::make_index_sequence< D >{} );
   auto ir = make_rule< Q >( std::make_index_sequence< D >{} );
   double sigma = -1.0, kappa = ( P + 1 ) * ( P + 1 );
   if ( std::sscanf( a.bc.c_str(), "s%lf_k%lf", &sigma, &kappa ) != 2 )
   {
      std::fprintf( stderr, "diffusion parameters must read s<sigma>_k<kappa>, got %s\n", a.bc.c_str() );
      return 2;
   }
   const std::string coef = a.coef;
   auto mu = [coef]( const std::array< Real, D > & X ) -> Real { return diffusivity< D >( coef, X ); };
   auto op = MakeDiffusionOperator< SerialKernelConfiguration >( fes, ir, mu, sigma, kappa );
   return finish( a, op, fes.GetNumberOfFiniteElementDofs(), fes.GetNumberOfFiniteElements() );
}

#define GD_DIFF( D, P, Q ) { D, P, Q, 'L', &run_diff< D, P, Q, true > }
#define GD_DIFF_G( D, P, Q ) { D, P, Q, 'G', &run_diff< D, P, Q, false > }

const CaseEntry kDiff[] = {
   GD_DIFF( 1, 1, 4 ),   GD_DIFF_G( 1, 3, 5 ), GD_DIFF_G( 2, 2, 4 ), GD_DIFF( 2, 3, 4 ),
   GD_DIFF( 3, 1, 4 ),   GD_DIFF( 3, 2, 5 ),   GD_DIFF( 3, 3, 6 ),   GD_DIFF( 3, 3, 4 ),
   GD_DIFF( 3, 3, 5 ),   GD_DIFF_G( 3, 2, 4 ), GD_DIFF_G( 6, 1, 3 ),
};

} // namespace

int run_diffusion_case( const Args & a )
{
   for ( const CaseEntry & e : kDiff )
      if ( e.D == a.D && e.P == a.P && e.Q == a.Q && e.basis == a.basis[0] ) return e.fn( a );
   std::fprintf( stderr, "diff case D=%d P=%d Q=%d basis=%s not instantiated\n", a.D, a.P, a.Q, a.basis.c_str() );
   return 2;
}

} // namespace gdref
