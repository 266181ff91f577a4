// Harness that instantiates GenDiL's own CPU operators (header-only C++20,
// SerialKernelConfiguration = OpenMP over elements) and writes input/output
// vectors for golden fixtures, or times the apply for the CPU baseline.
//
// This file is test infrastructure: it is compiled against the reference
// headers in place (see oracle/Makefile) into oracle/_ref/ and is never
// linked into the product library.  Shared pieces are in ref_common.hpp; the
// p-adaptive block composites are in ref_padapt.cpp.
//
// Usage:
//   gendil_ref fixture <op> <D> <P> <Q> <basis> <n> <periodic> <coef> <bc> <seed> <out.bin>
//   gendil_ref bench   <op> <D> <P> <Q> <basis> <n> <periodic> <coef> <bc> <seed> <pairs>
//   gendil_ref bicgstab|gmres <op> ... <out.bin>: GenDiL's BiCGSTAB_no_alloc or
//              GMRES_no_alloc (restart 30) on M + 0.05 A(coef, inflow bc)
//   gendil_ref cg <op> ... <out.bin>: GenDiL's ConjugateGradient on M(coef)
//              (solves: b = FillRandom(seed), x0 = 0, tol 1e-10, at most 500
//              iterations; output int64 n, b, x, iterations, success, residual)
//   op    : mass | adv | advface (face-centric operator on the interior faces)
//           | intermesh_<F>_<PP> (3D, GLL: face F of an n^3 minus mesh of order
//           P glued to an n^3 plus mesh of order PP, q shared; vector [minus|plus])
//           | hmesh_<F>_<R> (GLL order P: face F of a coarse n^D mesh glued to
//           a mesh R times finer along the face; vector [coarse|fine])
//           | padapt_<F>_<PP> (two GLL blocks of orders P and PP, each with its
//           own element-centric operator on q = p+1 points, coupled by the
//           intermesh face operator on face F with q = max(P, PP) + 1; see
//           ref_padapt.cpp)
//           | diff (SIPDG diffusion, ref_diffusion.cpp; coef = diffusivity
//           one | poly | trig, bc = s<sigma>_k<kappa>)
//   basis : L (Gauss-Lobatto nodes) | G (Gauss-Legendre nodes)
//   coef  : mass -> one | poly ;  adv -> c2 | bubble | const | trig | phase | yzx | ones
//   bc    : none (boundary faces skipped) | zero | poly
// fixture output: int64 ndofs, then x[ndofs] (FillRandom(seed)), then y[ndofs].

#include "ref_common.hpp"

using namespace gendil;
using namespace gdref;

namespace {

// GenDiL's nonconforming face operator (advectionoperator.hpp:578-676,
// 1210-1396) over CartesianIntermeshFaceConnectivity<3, F>: face F of an n^3
// minus mesh of GLL order PM at the origin glued to an n^3 plus mesh of order
// PP shifted by one unit along the face normal, Q points shared by both
// sides, the minus space first in the vector (as
// tests/integration/gendil/test-advection-p-adaptivity.cpp:13-100).
template < int PM, int PP, int Q, int F >
int run_intermesh( const Args & a )
{
   constexpr int D = 3;
   constexpr int axis = F % D;
   const Real h = 1.0 / a.n;
   Point< D > om = { 0.0, 0.0, 0.0 };
   Point< D > opl = om;
   opl[axis] = F >= D ? 1.0 : -1.0;
   Cartesian3DMesh mesh_m( h, h, h, a.n, a.n, a.n, om );
   Cartesian3DMesh mesh_p( h, h, h, a.n, a.n, a.n, opl );
   auto fe_m = MakeLobattoFiniteElement( FiniteElementOrders< PM, PM, PM >{} );
   auto fe_p = MakeLobattoFiniteElement( FiniteElementOrders< PP, PP, PP >{} );
   L2Restriction res_m{ 0 };
   auto fes_m = MakeFiniteElementSpace( mesh_m, fe_m, res_m );
   const Integer nm = fes_m.GetNumberOfFiniteElementDofs();
   L2Restriction res_p{ nm };
   auto fes_p = MakeFiniteElementSpace( mesh_p, fe_p, res_p );
   const Integer np = fes_p.GetNumberOfFiniteElementDofs();
   auto ir = MakeIntegrationRule( IntegrationRuleNumPoints< Q, Q, Q >{} );
   const std::array< GlobalIndex, D > sizes{ (GlobalIndex)a.n, (GlobalIndex)a.n, (GlobalIndex)a.n };
   CartesianIntermeshFaceConnectivity< D, F > conn( sizes, sizes );
   const std::string coef = a.coef;
   auto adv = [coef]( const std::array< Real, D > & X, Real ( &v )[D] ) { advection_field< D >( coef, X, v ); };
   auto op = MakeAdvectionFaceOperator< SerialKernelConfiguration >( fes_m, fes_p, conn, ir, adv );
   return finish( a, op, nm + np, fes_m.GetNumberOfFiniteElements() + fes_p.GetNumberOfFiniteElements() );
}

struct IntermeshEntry { int PM, PP, Q, F; Runner fn; };

// GenDiL's nonconforming face operator over
// NonconformingCartesianIntermeshFaceConnectivity<D, F>
// (nonconformingcartesianintermeshfaceconnectivity.hpp:73-243): face F of a
// coarse n^D mesh (minus, h = 1/n, at the origin) glued to a fine mesh (plus)
// with R times the cells along every tangential axis (h = 1/(R n) there),
// shifted by one unit along the face normal, GLL order P on both sides, Q
// points shared, vector [coarse | fine] (as
// tests/integration/gendil/test-advection-h-adaptivity-2d.cpp:66-140).
template < int D, int P, int Q, int F, int R >
int run_hmesh( const Args & a )
{
   constexpr int axis = F % D;
   std::array< GlobalIndex, D > sc, sf;
   std::array< Real, D > hc, hf;
   Point< D > oc, of;
   for ( int d = 0; d < D; ++d )
   {
      sc[d] = a.n;
      sf[d] = d == axis ? a.n : R * a.n;
      hc[d] = 1.0 / a.n;
      hf[d] = d == axis ? 1.0 / a.n : 1.0 / ( R * a.n );
      oc[d] = 0.0;
      of[d] = d == axis ? ( F >= D ? 1.0 : -1.0 ) : 0.0;
   }
   CartesianMesh< D > mesh_c( sc, hc, oc, false );
   CartesianMesh< D > mesh_f( sf, hf, of, false );
   auto fe = make_gll< P >( std::make_index_sequence< D >{} );
   L2Restriction res_c{ 0 };
   auto fes_c = MakeFiniteElementSpace( mesh_c, fe, res_c );
   const Integer nc = fes_c.GetNumberOfFiniteElementDofs();
   L2Restriction res_f{ nc };
   auto fes_f = MakeFiniteElementSpace( mesh_f, fe, res_f );
   const Integer nf = fes_f.GetNumberOfFiniteElementDofs();
   auto ir = make_rule< Q >( std::make_index_sequence< D >{} );
   NonconformingCartesianIntermeshFaceConnectivity< D, F > conn( sc, sf );
   const std::string coef = a.coef;
   auto adv = [coef]( const std::array< Real, D > & X, Real ( &v )[D] ) { advection_field< D >( coef, X, v ); };
   auto op = MakeAdvectionFaceOperator< SerialKernelConfiguration >( fes_c, fes_f, conn, ir, adv );
   return finish( a, op, nc + nf, fes_c.GetNumberOfFiniteElements() + fes_f.GetNumberOfFiniteElements() );
}

struct HmeshEntry { int D, P, Q, F, R; Runner fn; };


// AdvectionFaceOperator over CartesianInteriorFaceConnectivity
// (advectionoperator.hpp:1105-1180, cartesianglobalfaceconnectivity.hpp:15-95):
// every interior face once, y accumulated (+ minus side, - plus side)
template < int D, int P, int Q, bool Lobatto >
int run_face( const Args & a )
{
   auto fes = make_space< D, P, Q, Lobatto >( a, std::make_index_sequence< D >{} );
   auto ir = make_rule< Q >( std::make_index_sequence< D >{} );
   const std::string coef = a.coef;
   auto adv = [coef]( const std::array< Real, D > & X, Real ( &v )[D] ) { advection_field< D >( coef, X, v ); };
   std::array< GlobalIndex, D > sizes;
   for ( int d = 0; d < D; ++d ) sizes[d] = a.n;
   auto conn = MakeCartesianInteriorFaceConnectivity< D >( sizes );
   auto op = MakeAdvectionFaceOperator< SerialKernelConfiguration >( fes, conn, ir, adv );
   return finish( a, op, fes.GetNumberOfFiniteElementDofs(), fes.GetNumberOfFiniteElements() );
}

// The operator of GenDiL's 6D benchmark (benchmarks/benchmark-advection-6d.cpp:
// 14-98): a product of a 3D (h = 1), a 2D (h = 1.2345) and a 1D (h = 1)
// Cartesian mesh with n elements per axis, Legendre p = 1 elements, 3 Gauss
// points per axis, beta = 1 in every direction, no boundary data.
int run_product( const Args & a )
{
   Cartesian3DMesh mesh_1( 1.0, a.n, a.n, a.n );
   Cartesian2DMesh mesh_2( 1.2345, a.n, a.n );
   Cartesian1DMesh mesh_3( 1.0, a.n );
   auto mesh = MakeCartesianProductMesh( mesh_1, mesh_2, mesh_3 );
   FiniteElementOrders< 1, 1, 1, 1, 1, 1 > orders;
   auto fe = MakeLegendreFiniteElement( orders );
   auto fes = MakeFiniteElementSpace( mesh, fe );
   IntegrationRuleNumPoints< 3, 3, 3, 3, 3, 3 > nq;
   auto ir = MakeIntegrationRule( nq );
   auto adv = []( const std::array< Real, 6 > &, Real ( &v )[6] )
   {
      for ( int d = 0; d < 6; ++d ) v[d] = 1.0;
   };
   auto op = MakeAdvectionOperator< SerialKernelConfiguration >( fes, ir, adv );
   return finish( a, op, fes.GetNumberOfFiniteElementDofs(), fes.GetNumberOfFiniteElements() );
}

const CaseEntry kEntries[] = {
   GD_CASE( 3, 0, 2 ), GD_CASE( 3, 1, 2 ), GD_CASE( 3, 1, 3 ), GD_CASE( 3, 2, 3 ),
   GD_CASE( 3, 2, 4 ), GD_CASE( 3, 2, 5 ), GD_CASE( 3, 3, 4 ), GD_CASE( 3, 3, 5 ),
   GD_CASE( 3, 3, 6 ), GD_CASE( 3, 4, 5 ), GD_CASE( 3, 4, 6 ),
   GD_CASE_G( 3, 3, 4 ), GD_CASE_G( 3, 2, 3 ),
   GD_CASE( 1, 3, 4 ), GD_CASE( 2, 2, 3 ), GD_CASE( 2, 3, 4 ),
   GD_CASE( 4, 1, 2 ), GD_CASE( 4, 2, 3 ),
   GD_CASE( 6, 1, 2 ), GD_CASE( 6, 2, 3 ),
};

const CaseEntry kFaceEntries[] = {
   { 3, 3, 4, 'L', &run_face< 3, 3, 4, true > }, { 3, 2, 3, 'G', &run_face< 3, 2, 3, false > },
   { 2, 2, 3, 'L', &run_face< 2, 2, 3, true > }, { 6, 1, 2, 'L', &run_face< 6, 1, 2, true > },
};

const HmeshEntry kHmesh[] = {
   { 2, 2, 4, 2, 2, &run_hmesh< 2, 2, 4, 2, 2 > }, { 3, 2, 3, 3, 2, &run_hmesh< 3, 2, 3, 3, 2 > },
   { 3, 1, 2, 0, 2, &run_hmesh< 3, 1, 2, 0, 2 > }, { 3, 3, 4, 4, 3, &run_hmesh< 3, 3, 4, 4, 3 > },
};

const IntermeshEntry kIntermesh[] = {
   { 1, 3, 5, 3, &run_intermesh< 1, 3, 5, 3 > }, { 3, 1, 5, 0, &run_intermesh< 3, 1, 5, 0 > },
   { 2, 2, 4, 0, &run_intermesh< 2, 2, 4, 0 > }, { 2, 3, 4, 4, &run_intermesh< 2, 3, 4, 4 > },
   { 3, 2, 4, 2, &run_intermesh< 3, 2, 4, 2 > },
};

} // namespace

int main( int argc, char ** argv )
{
   if ( argc < 13 )
   {
      std::fprintf( stderr, "usage: %s fixture|bicgstab|bench op D P Q basis n periodic coef bc seed out|pairs\n", argv[0] );
      return 2;
   }
   Args a;
   a.mode = argv[1];
   a.op = argv[2];
   a.D = std::atoi( argv[3] );
   a.P = std::atoi( argv[4] );
   a.Q = std::atoi( argv[5] );
   a.basis = argv[6];
   a.n = std::atoi( argv[7] );
   a.periodic = std::atoi( argv[8] );
   a.coef = argv[9];
   a.bc = argv[10];
   a.seed = (unsigned)std::strtoul( argv[11], nullptr, 10 );
   if ( a.mode == "fixture" || a.mode == "bicgstab" || a.mode == "gmres" || a.mode == "cg" ) a.out = argv[12];
   else a.pairs = std::atoi( argv[12] );
   if ( a.op == "adv6dprod" ) return run_product( a );
   if ( a.op.rfind( "padapt_", 0 ) == 0 ) return run_padapt_case( a );
   if ( a.op == "diff" ) return run_diffusion_case( a );
   if ( a.op == "advface" )
   {
      for ( const CaseEntry & e : kFaceEntries )
         if ( e.D == a.D && e.P == a.P && e.Q == a.Q && e.basis == a.basis[0] ) return e.fn( a );
      std::fprintf( stderr, "advface case D=%d P=%d Q=%d basis=%s not instantiated\n", a.D, a.P, a.Q, a.basis.c_str() );
      return 2;
   }
   if ( a.op.rfind( "hmesh_", 0 ) == 0 )
   {
      int F = -1, R = -1;
      if ( std::sscanf( a.op.c_str(), "hmesh_%d_%d", &F, &R ) != 2 ) return 2;
      for ( const HmeshEntry & e : kHmesh )
         if ( e.D == a.D && e.P == a.P && e.Q == a.Q && e.F == F && e.R == R ) return e.fn( a );
      std::fprintf( stderr, "hmesh case D=%d F=%d P=%d Q=%d R=%d not instantiated\n", a.D, F, a.P, a.Q, R );
      return 2;
   }
   if ( a.op.rfind( "intermesh_", 0 ) == 0 )
   {
      int F = -1, PP = -1;
      if ( std::sscanf( a.op.c_str(), "intermesh_%d_%d", &F, &PP ) != 2 ) return 2;
      for ( const IntermeshEntry & e : kIntermesh )
         if ( e.PM == a.P && e.PP == PP && e.Q == a.Q && e.F == F ) return e.fn( a );
      std::fprintf( stderr, "intermesh case F=%d PM=%d PP=%d Q=%d not instantiated\n", F, a.P, PP, a.Q );
      return 2;
   }
   for ( const CaseEntry & e : kEntries )
   {
      if ( e.D == a.D && e.P == a.P && e.Q == a.Q && e.basis == a.basis[0] ) return e.fn( a );
   }
   return run_highdim_case( a );
}
