// Cases of the GenDiL harness (test infrastructure; see ref_harness.cpp) at
// the (order, points) pairs of GenDiL's range benchmarks in 3D, 4D and 6D with
// Legendre elements (benchmarks/range-benchmark-advection-{3,4,6}d.cpp,
// range-benchmark-mass-{3,4,6}d.cpp): 3D q = p + 2 up to p = 6, 4D up to
// p = 4, 6D up to p = 2 (advection q = p + 1, mass q = p + 2 on the device).
#include "ref_common.hpp"

namespace gdref {

namespace {

const CaseEntry kEntriesRange[] = {
   GD_CASE_G( 3, 5, 7 ), GD_CASE_G( 3, 6, 8 ), GD_CASE_G( 4, 4, 5 ), GD_CASE_G( 4, 4, 6 ),
   GD_CASE_G( 6, 2, 3 ), GD_CASE_G( 6, 2, 4 ),
};

} // namespace

int run_rangebench_case( const Args & a )
{
   for ( const CaseEntry & e : kEntriesRange )
      if ( e.D == a.D && e.P == a.P && e.Q == a.Q && e.basis == a.basis[0] ) return e.fn( a );
   std::fprintf( stderr, "case D=%d P=%d Q=%d basis=%s not instantiated\n", a.D, a.P, a.Q, a.basis.c_str() );
   return 2;
}

} // namespace gdref
