// momenta.h -- the reference's header name for momenta; the class lives in lattice.h.
#ifndef JW_HOST_MOMENTA_H
#define JW_HOST_MOMENTA_H
#include "lattice.h"
#endif
