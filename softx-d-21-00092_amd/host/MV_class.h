// MV_class.h -- the reference's header name for MV_class; the class lives in lattice.h.
#ifndef JW_HOST_MV_CLASS_H
#define JW_HOST_MV_CLASS_H
#include "lattice.h"
#endif
