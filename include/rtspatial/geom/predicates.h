// Predicate (reference include/rtspatial/geom/predicates.h:7-10) plus kWithin:
// a data box lying inside the query envelope.
#ifndef RTSPATIAL_AMD_GEOM_PREDICATES_H
#define RTSPATIAL_AMD_GEOM_PREDICATES_H

#include "rts_c.h"

namespace rtspatial {

enum class Predicate {
  kIntersects = RTS_INTERSECTS,
  kContains = RTS_CONTAINS,
  kWithin = RTS_WITHIN,
};

}  // namespace rtspatial
#endif
