// Selects the collecting handler for SpatialIndex::Query(..., void* arg, ...):
// `arg` is the device rts_dev_queue of a Queue (Queue::DeviceObject()), to
// which (geom_id, query_id) pairs are appended (reference
// include/rtspatial/default_handlers/collecting_handlers.h:9-31). This is the
// default when no handler header is included.
#ifndef RTSPATIAL_AMD_DEFAULT_HANDLERS_COLLECTING_H
#define RTSPATIAL_AMD_DEFAULT_HANDLERS_COLLECTING_H
#include "rts_c.h"
#ifndef RTSPATIAL_HANDLER
#define RTSPATIAL_HANDLER RTS_HANDLER_COLLECT
#endif
#endif
