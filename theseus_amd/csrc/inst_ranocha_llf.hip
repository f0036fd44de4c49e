// Kernel instantiations: Ranocha volume + Lax-Friedrichs face (3-D Taylor-Green
// workload), plus the tile-shape variants the serving pool tunes over and
// the kernel that advances pools of orders 2, 3 and 4 in one launch.
#include "registry.hpp"

namespace thx {
namespace {
template <int DIM>
struct Orders {
  FluxRegistrar<DIM, 2, 5, 1> r1;
  FluxRegistrar<DIM, 2, 5, 2> r2;
  FluxRegistrar<DIM, 2, 5, 3> r3;
  FluxRegistrar<DIM, 2, 5, 4> r4;
  FluxRegistrar<DIM, 2, 5, 5> r5;
};
Orders<2> orders2;
Orders<3> orders3;
// tile-shape variants for the serving pool (3-D, the orders it serves)
TileRegistrar<3, 2, 5, 2> tiles2;
TileRegistrar<3, 2, 5, 3> tiles3;
TileRegistrar<3, 2, 5, 4> tiles4;
// orders 2, 3 and 4 in one launch
OrdersRegistrar<2, 5> orders234;
}  // namespace
}  // namespace thx
