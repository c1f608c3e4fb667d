/* Target contexts linked into this library. */
#ifndef MUBLIS_TARGET_REGISTRY_H
#define MUBLIS_TARGET_REGISTRY_H

#include "mublis_instantiate.h"

extern const mublis_context_t gfx950_context;

#endif
