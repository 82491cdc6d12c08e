// positions.h -- the reference's header name for positions; the class lives in lattice.h.
#ifndef JW_HOST_POSITIONS_H
#define JW_HOST_POSITIONS_H
#include "lattice.h"
#endif
