// rand_class.h -- the reference's header name for rand_class; the class lives in lattice.h.
#ifndef JW_HOST_RAND_CLASS_H
#define JW_HOST_RAND_CLASS_H
#include "lattice.h"
#endif
