// Kernel instantiations independent of the flux pairing: Navier-Stokes
// pass 1 (BR1 gradient + viscous flux) and the stability reduction.
#include "registry.hpp"

namespace thx {
namespace {
template <int DIM>
struct Orders {
  ShapeRegistrar<DIM, 1> r1;
  ShapeRegistrar<DIM, 2> r2;
  ShapeRegistrar<DIM, 3> r3;
  ShapeRegistrar<DIM, 4> r4;
  ShapeRegistrar<DIM, 5> r5;
};
Orders<2> orders2;
Orders<3> orders3;
}  // namespace
}  // namespace thx
