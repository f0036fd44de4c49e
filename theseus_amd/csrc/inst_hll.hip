// Kernel instantiations: reference 'hll': central volume + HLL face.
#include "registry.hpp"

namespace thx {
namespace {
template <int DIM>
struct Orders {
  FluxRegistrar<DIM, 0, 4, 1> r1;
  FluxRegistrar<DIM, 0, 4, 2> r2;
  FluxRegistrar<DIM, 0, 4, 3> r3;
  FluxRegistrar<DIM, 0, 4, 4> r4;
  FluxRegistrar<DIM, 0, 4, 5> r5;
};
Orders<2> orders2;
Orders<3> orders3;
}  // namespace
}  // namespace thx
