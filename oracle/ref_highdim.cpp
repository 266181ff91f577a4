// 5D cases of the GenDiL harness (test infrastructure; see ref_harness.cpp):
// GenDiL's mass and advection operators on CartesianMesh<5>, among them the
// (order, points) pairs of its 5D range benchmarks
// (benchmarks/range-benchmark-advection-5d.cpp, range-benchmark-mass-5d.cpp).
#include "ref_common.hpp"

namespace gdref {

namespace {

const CaseEntry kEntries5[] = {
   GD_CASE_G( 5, 1, 2 ), GD_CASE( 5, 2, 3 ), GD_CASE( 5, 3, 4 ), GD_CASE_G( 5, 2, 4 ),
};

} // namespace

int run_highdim_case( const Args & a )
{
   for ( const CaseEntry & e : kEntries5 )
      if ( e.D == a.D && e.P == a.P && e.Q == a.Q && e.basis == a.basis[0] ) return e.fn( a );
   return run_rangebench_case( a );
}

} // namespace gdref
