// mpi_fftw_class.h -- the reference's header name for fftw2D; the class lives in field.h.
#ifndef JW_HOST_MPI_FFTW_CLASS_H
#define JW_HOST_MPI_FFTW_CLASS_H
#include "field.h"
#endif
