// Kernel instantiations: reference 'ranocha': EC volume + EC/Roe-dissipation face.
#include "registry.hpp"

namespace thx {
namespace {
template <int DIM>
struct Orders {
  FluxRegistrar<DIM, 2, 1, 1> r1;
  FluxRegistrar<DIM, 2, 1, 2> r2;
  FluxRegistrar<DIM, 2, 1, 3> r3;
  FluxRegistrar<DIM, 2, 1, 4> r4;
  FluxRegistrar<DIM, 2, 1, 5> r5;
};
Orders<2> orders2;
Orders<3> orders3;
}  // namespace
}  // namespace thx
