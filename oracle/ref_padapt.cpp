// GenDiL's p-adaptivity, run by GenDiL's own operators (test infrastructure,
// compiled against the reference headers into oracle/_ref/, never linked into
// the product library).
//
// GenDiL has no per-element order.  Its p-adaptive discretisation is a set of
// Cartesian blocks of uniform compile-time order, one vector holding every
// block at its L2Restriction shift, each block applied by its own
// element-centric advection operator (which skips the block's boundary faces:
// MakeAdvectionOperator without boundary data, faceloop.hpp:66-72), and the
// interfaces between blocks applied by the nonconforming face operator over
// CartesianIntermeshFaceConnectivity with one shared rule for both sides
// (advectionoperator.hpp:596-676, 1210-1396; the pattern of
// tests/integration/gendil/test-advection-p-adaptivity.cpp:13-100).
//
// Case padapt_<F>_<PP>, dimension D, order P (= PM):
//   minus block: GLL order PM, n cells along every axis, origin 0;
//   plus block:  GLL order PP, n + 1 cells along the interface axis F mod D
//                (n along the others), on the far side of face F of the minus
//                block (origin +1 along the axis for F >= D, -(n+1)/n for F < D);
//   both blocks: h = 1/n on every axis; periodic along every tangential axis
//                when `periodic` is 1, never along the interface axis;
//   cell rules:  q = p + 1 Gauss points per block; interface rule q = max + 1;
//   y = A_minus x + A_plus x + F x  (the block operators overwrite their part
//   of y, the face operator accumulates), x = FillRandom(seed) over
//   [minus | plus].
// This is the same operator as one Cartesian mesh with per-element orders and
// the face rule q = max(q-, q+) on every face.

#include "ref_common.hpp"

namespace gdref {

namespace {

template < int D, int PM, int PP, int F >
int run_padapt( const Args & a )
{
   constexpr int axis = F % D;
   constexpr int QM = PM + 1;
   constexpr int QP = PP + 1;
   constexpr int QF = ( PM > PP ? PM : PP ) + 1;
   std::array< GlobalIndex, D > sm, sp;
   std::array< Real, D > h;
   Point< D > om, opl;
   std::array< int, D > per;
   for ( int d = 0; d < D; ++d )
   {
      sm[d] = a.n;
      sp[d] = d == axis ? a.n + 1 : a.n;
      h[d] = 1.0 / a.n;
      om[d] = 0.0;
      opl[d] = d != axis ? 0.0 : ( F >= D ? 1.0 : -( a.n + 1 ) * h[d] );
      per[d] = d == axis ? 0 : ( a.periodic != 0 );
   }
   CartesianMesh< D > mesh_m( sm, h, om, per );
   CartesianMesh< D > mesh_p( sp, h, opl, per );
   auto fe_m = make_gll< PM >( std::make_index_sequence< D >{} );
   auto fe_p = make_gll< PP >( std::make_index_sequence< D >{} );
   L2Restriction res_m{ 0 };
   auto fes_m = MakeFiniteElementSpace( mesh_m, fe_m, res_m );
   const Integer nm = fes_m.GetNumberOfFiniteElementDofs();
   L2Restriction res_p{ nm };
   auto fes_p = MakeFiniteElementSpace( mesh_p, fe_p, res_p );
   const Integer np = fes_p.GetNumberOfFiniteElementDofs();
   auto ir_m = make_rule< QM >( std::make_index_sequence< D >{} );
   auto ir_p = make_rule< QP >( std::make_index_sequence< D >{} );
   auto ir_f = make_rule< QF >( std::make_index_sequence< D >{} );
   const std::string coef = a.coef;
   auto adv = [coef]( const std::array< Real, D > & X, Real ( &v )[D] ) { advection_field< D >( coef, X, v ); };
   auto A_m = MakeAdvectionOperator< SerialKernelConfiguration >( fes_m, ir_m, adv );
   auto A_p = MakeAdvectionOperator< SerialKernelConfiguration >( fes_p, ir_p, adv );
   CartesianIntermeshFaceConnectivity< D, F > conn( sm, sp );
   auto A_f = MakeAdvectionFaceOperator< SerialKernelConfiguration >( fes_m, fes_p, conn, ir_f, adv );
   auto op = [&]( const Vector & x, Vector & y )
   {
      A_m( x, y );
      A_p( x, y );
      A_f( x, y );
   };
   return finish( a, op, nm + np, fes_m.GetNumberOfFiniteElements() + fes_p.GetNumberOfFiniteElements() );
}

struct PadaptEntry { int D, PM, PP, F; Runner fn; };

#define GD_PA( D, PM, PP, F ) { D, PM, PP, F, &run_padapt< D, PM, PP, F > }

const PadaptEntry kPadapt[] = {
   // +x interface: every ordered pair of distinct orders 1..3
   GD_PA( 3, 1, 2, 3 ), GD_PA( 3, 2, 1, 3 ), GD_PA( 3, 1, 3, 3 ),
   GD_PA( 3, 3, 1, 3 ), GD_PA( 3, 2, 3, 3 ), GD_PA( 3, 3, 2, 3 ),
   // -y interface: every ordered pair
   GD_PA( 3, 1, 2, 1 ), GD_PA( 3, 2, 1, 1 ), GD_PA( 3, 1, 3, 1 ),
   GD_PA( 3, 3, 1, 1 ), GD_PA( 3, 2, 3, 1 ), GD_PA( 3, 3, 2, 1 ),
   // the other orientation of each axis, and the z axis
   GD_PA( 3, 1, 3, 0 ), GD_PA( 3, 3, 2, 0 ), GD_PA( 3, 2, 1, 4 ),
   GD_PA( 3, 3, 1, 4 ), GD_PA( 3, 2, 3, 5 ), GD_PA( 3, 1, 2, 2 ),
   // phase space: blocks split along a velocity axis and along a space axis
   GD_PA( 6, 1, 2, 9 ), GD_PA( 6, 2, 1, 0 ),
};

} // namespace

int run_padapt_case( const Args & a )
{
   int F = -1, PP = -1;
   if ( std::sscanf( a.op.c_str(), "padapt_%d_%d", &F, &PP ) != 2 ) return 2;
   if ( a.basis != "L" )
   {
      std::fprintf( stderr, "padapt cases use the GLL basis\n" );
      return 2;
   }
   for ( const PadaptEntry & e : kPadapt )
      if ( e.D == a.D && e.PM == a.P && e.PP == PP && e.F == F ) return e.fn( a );
   std::fprintf( stderr, "padapt case D=%d F=%d PM=%d PP=%d not instantiated\n", a.D, F, a.P, PP );
   return 2;
}

} // namespace gdref
