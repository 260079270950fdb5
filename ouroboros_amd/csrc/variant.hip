// One MultiOuro* instantiation per translation unit; OURO_VARIANT selects which (0..5, the
// OURO_PQ ... OURO_VLCQ order of ouro_capi.h).
#include "manager.hpp"

#ifndef OURO_VARIANT
#error "compile with -DOURO_VARIANT=<0..5>"
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
#endif
}  // namespace ouro_lib
