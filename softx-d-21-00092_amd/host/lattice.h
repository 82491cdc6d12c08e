// lattice.h -- the small classes around the fields, with the reference's names
// and accessors: mpi_class (single rank), momenta, positions, MV_class,
// rand_class.  One GPU holds the whole lattice, so the MPI grid is 1 x 1.
#ifndef JW_HOST_LATTICE_H
#define JW_HOST_LATTICE_H

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include <vector>

#include "config.h"

// mpi_class (mpi_class.h:38-167): the reference takes the process grid from
// argv[1], argv[2] (mpi_class.h:80-81) and aborts with exit code 0 when it does
// not match the ranks (mpi_init.cpp:47-119).  Here the grid must be 1 x 1; any
// other grid is an error with a non-zero exit code.
class mpi_class {
    int proc_grid[2] = {1, 1};

   public:
    mpi_class(int argc, char* argv[]) {
        if (argc > 2) {
            proc_grid[0] = atoi(argv[1]);
            proc_grid[1] = atoi(argv[2]);
        }
    }
    mpi_class() {}
    int mpi_init(config* cnfg) {
        if (proc_grid[0] != 1 || proc_grid[1] != 1) {
            fprintf(stderr, "mpi_class: process grid %d x %d requested; this solver runs one lattice per GPU (1 x 1)\n",
                    proc_grid[0], proc_grid[1]);
            exit(2);
        }
        cnfg->Nxl = cnfg->Nx;
        cnfg->Nyl = cnfg->Ny;
        cnfg->Nxl_buf = cnfg->Nx;
        cnfg->Nyl_buf = cnfg->Ny;
        return 1;
    }
    int mpi_exchange_grid() { return 1; }
    int getRank() { return 0; }
    int getSize() { return 1; }
    int getPosX() { return 0; }
    int getPosY() { return 0; }
};

// momenta (momenta.h:39-140): lattice momenta of the (single) local block
class momenta {
    std::vector<double> tbl_phat2, tbl_pbarx, tbl_pbary;
    int Nxl, Nyl;

   public:
    momenta(config* cnfg, mpi_class*) : Nxl(cnfg->Nxl), Nyl(cnfg->Nyl) {}
    int set() {
        tbl_phat2.resize((size_t)Nxl * Nyl);
        tbl_pbarx.resize((size_t)Nxl * Nyl);
        tbl_pbary.resize((size_t)Nxl * Nyl);
        for (int i = 0; i < Nxl; i++)
            for (int j = 0; j < Nyl; j++) {
                const double sx = sin(M_PI * i / (1.0 * Nxl)), sy = sin(M_PI * j / (1.0 * Nyl));
                tbl_phat2[(size_t)i * Nyl + j] = 4.0 * pow(sx, 2.0) + 4.0 * pow(sy, 2.0);
                tbl_pbarx[(size_t)i * Nyl + j] = sin(2.0 * M_PI * i / (1.0 * Nxl));
                tbl_pbary[(size_t)i * Nyl + j] = sin(2.0 * M_PI * j / (1.0 * Nyl));
            }
        return 1;
    }
    double phat2(int i) { return tbl_phat2[i]; }
    double pbarX(int i) { return tbl_pbarx[i]; }
    double pbarY(int i) { return tbl_pbary[i]; }
};

// positions (positions.h:38-180): position-space kernel tables.  Only the
// position-space evolution uses them; kept for the class surface.
class positions {
    std::vector<double> tbl_xhatx, tbl_xhaty, tbl_xbar2;
    int Nxg, Nyg;

   public:
    positions(config* cnfg, mpi_class*) : Nxg(cnfg->Nx), Nyg(cnfg->Ny) {}
    int set() {
        tbl_xhatx.resize((size_t)Nxg * Nyg);
        tbl_xhaty.resize((size_t)Nxg * Nyg);
        tbl_xbar2.resize((size_t)Nxg * Nyg);
        for (int i = 0; i < Nxg; i++)
            for (int j = 0; j < Nyg; j++) {
                tbl_xhatx[(size_t)i * Nyg + j] = 0.5 * Nxg * sin(2.0 * M_PI * i / (1.0 * Nxg)) / M_PI;
                tbl_xhaty[(size_t)i * Nyg + j] = 0.5 * Nyg * sin(2.0 * M_PI * j / (1.0 * Nyg)) / M_PI;
                const double bx = Nxg * sin(M_PI * i / (1.0 * Nxg)) / M_PI, by = Nyg * sin(M_PI * j / (1.0 * Nyg)) / M_PI;
                tbl_xbar2[(size_t)i * Nyg + j] = pow(bx, 2.0) + pow(by, 2.0);
            }
        return 1;
    }
    double xhatX(int i) { return tbl_xhatx[i]; }
    double xhatY(int i) { return tbl_xhaty[i]; }
    double xbar2(int i) { return tbl_xbar2[i]; }
};

// MV_class (MV_class.h:46-60)
class MV_class {
   public:
    double g_parameter;
    double mu_parameter;
    int Ny_parameter;
    MV_class(double g, double mu, int N) : g_parameter(g), mu_parameter(mu), Ny_parameter(N) {}
};

// rand_class (rand_class.h:47-68): uniform numbers in (0, 1).  The reference
// wraps ranlux24_base seeded with seed + 64 rank + thread; this one draws from
// the Philox4x32-10 stream with key {seed, 0xFFFFFFFF} and a running counter,
// two 32-bit words per 53-bit uniform.
class rand_class {
    uint32_t key0, key1 = 0xFFFFFFFFu;
    uint64_t counter = 0;
    uint32_t buf[4];
    int used = 4;
    void refill() {
        uint32_t c0 = (uint32_t)counter, c1 = (uint32_t)(counter >> 32), c2 = 0, c3 = 0;
        uint32_t k0 = key0, k1 = key1;
        for (int r = 0; r < 10; r++) {
            const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
            const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
            c0 = n0; c1 = (uint32_t)p1; c2 = n2; c3 = (uint32_t)p0;
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        buf[0] = c0; buf[1] = c1; buf[2] = c2; buf[3] = c3;
        used = 0;
        counter++;
    }

   public:
    rand_class(mpi_class* mpi, config* cnfg) : key0(cnfg->seed + 64u * (uint32_t)mpi->getRank()) {}
    double get() {
        if (used > 2) refill();
        const uint64_t w = ((uint64_t)buf[used] << 32) | buf[used + 1];
        used += 2;
        return ((double)(w >> 11) + 0.5) * 0x1.0p-53;
    }
};

#endif
