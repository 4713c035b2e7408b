// Umbrella header (reference include/rtspatial/rtspatial.h:1-11).
#ifndef RTSPATIAL_AMD_RTSPATIAL_H
#define RTSPATIAL_AMD_RTSPATIAL_H
#include "rtspatial/spatial_index.h"
#include "rtspatial/utils/memory.h"
#include "rtspatial/utils/shared_value.h"
#include "rtspatial/utils/stream.h"
#endif
