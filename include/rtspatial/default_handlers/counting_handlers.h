// Selects the counting handler for SpatialIndex::Query(..., void* arg, ...):
// `arg` is a device unsigned long long* to which the number of results is
// added (reference include/rtspatial/default_handlers/counting_handlers.h:7-25,
// which the reference injects into its shaders at build time). Include before
// rtspatial/spatial_index.h.
#ifndef RTSPATIAL_AMD_DEFAULT_HANDLERS_COUNTING_H
#define RTSPATIAL_AMD_DEFAULT_HANDLERS_COUNTING_H
#include "rts_c.h"
#ifndef RTSPATIAL_HANDLER
#define RTSPATIAL_HANDLER RTS_HANDLER_COUNT
#endif
#endif
