// One instantiation per translation unit; OURO_VARIANT selects which: 0..5 the MultiOuro* chains in
// the OURO_PQ ... OURO_VLCQ order of ouro_capi.h, 6.. the tuner's pool geometries.
#include "manager.hpp"

#ifndef OURO_VARIANT
#error "compile with -DOURO_VARIANT=<0..11>"
#endif

namespace ouro_lib {
#if OURO_VARIANT == 0
IManager* make_pq() { return new Manager<ouro::MultiOuroPQ>(OURO_PQ); }
#elif OURO_VARIANT == 1
IManager* make_cq() { return new Manager<ouro::MultiOuroCQ>(OURO_CQ); }
#elif OURO_VARIANT == 2
IManager* make_vapq() { return new Manager<ouro::MultiOuroVAPQ>(OURO_VAPQ); }
#elif OURO_VARIANT == 3
IManager* make_vacq() { return new Manager<ouro::MultiOuroVACQ>(OURO_VACQ); }
#elif OURO_VARIANT == 4
IManager* make_vlpq() { return new Manager<ouro::MultiOuroVLPQ>(OURO_VLPQ); }
#elif OURO_VARIANT == 5
IManager* make_vlcq() { return new Manager<ouro::MultiOuroVLCQ>(OURO_VLCQ); }
#elif OURO_VARIANT == 6
IManager* make_geo6() { return new Manager<ouro::OuroGeometry<ouro::PageQueueVA, 8192, 10>>(6); }
#elif OURO_VARIANT == 7
IManager* make_geo7() { return new Manager<ouro::OuroGeometry<ouro::PageQueueVA, 8192, 8>>(7); }
#elif OURO_VARIANT == 8
IManager* make_geo8() { return new Manager<ouro::OuroGeometry<ouro::PageQueueVA, 16384, 10>>(8); }
#elif OURO_VARIANT == 9
IManager* make_geo9() { return new Manager<ouro::OuroGeometry<ouro::PageQueueVA, 16384, 8>>(9); }
#elif OURO_VARIANT == 10
IManager* make_geo10() { return new Manager<ouro::MultiOuroCapacity<ouro::PageQueueVA, 4096>>(10); }
#elif OURO_VARIANT == 11
IManager* make_geo11() { return new Manager<ouro::MultiOuroCapacity<ouro::PageQueue, 1u << 20>>(11); }
#endif
}  // namespace ouro_lib
