// main_optimized.cpp -- production driver, the structure of the reference's
// main_optimized.cpp:64-379 on one GPU.
//
//   ./main_optimized 1 1 configfile        (process grid must be 1 x 1)
//
// Per stat sample: MV initial condition (main_optimized.cpp:178-191), then
// langevin_steps momentum-space Langevin steps (234-311) with the measurement
// every langevin_steps/measurements steps (317-336), then print (351-355).
// The Langevin step runs as the reference's per-function sequence (engine =
// reference, one step per call), as the library's generator-basis engine
// (engine = fused, the default; the steps between two measurements in one
// call), or with all stat samples evolving together in one batch (engine =
// batched).  precision = mixed selects the mixed-precision form for the
// latter two.  Noise: Philox stream {seed, stat sample}, the Langevin steps
// continuing the MV sheets' counter, so every engine draws the same noise.
//
// Deviations from the reference, all stated: sum/err are sized by
// `measurements` (the reference sizes them by langevin_steps,
// main_optimized.cpp:134-135, 134 GB at N = 2048); the output file root is
// config file_name (the reference ignores it and writes "output_optimized");
// POSITION_EVOLUTION and HATTA_COUPLING_CONSTANT are rejected.
#include <stdio.h>
#include <time.h>

#include <iostream>
#include <vector>

#include "config.h"
#include "field.h"
#include "lattice.h"

static double seconds_since(const timespec& t0) {
    timespec t1;
    clock_gettime(CLOCK_MONOTONIC, &t1);
    return (t1.tv_sec - t0.tv_sec) + (t1.tv_nsec - t0.tv_nsec) / 1e9;
}

int main(int argc, char* argv[]) {
    printf("INITIALIZATION\n");
    config* cnfg = new config;
    if (argc > 3 && !cnfg->read_config_from_file(argv[3])) return 1;
    if (cnfg->EvolutionChoice != MOMENTUM_EVOLUTION || cnfg->CouplingChoice == HATTA_COUPLING_CONSTANT) {
        fprintf(stderr, "only MOMENTUM_EVOLUTION with SQRT, NOISE, FIXED or CORRELATED coupling is supported\n");
        return 2;
    }
    mpi_class* mpi = new mpi_class(argc, argv);
    mpi->mpi_init(cnfg);
    mpi->mpi_exchange_grid();
    const int Nx = cnfg->Nx, Ny = cnfg->Ny;

    momenta* momtable = new momenta(cnfg, mpi);
    momtable->set();
    positions postable(cnfg, mpi);
    postable.set();
    MV_class* MVmodel = new MV_class(1.0, cnfg->mu / Nx, cnfg->elementaryWilsonLines);
    fftw2D* fourier2 = new fftw2D(cnfg);
    fourier2->init2D();

    printf("ALLOCATION\n");
    {
        lfield<double, 9> f(cnfg->Nxl, cnfg->Nyl);
        lfield<double, 9> uf(cnfg->Nxl, cnfg->Nyl);
        lfield<double, 9> xi_local_x(cnfg->Nxl, cnfg->Nyl);
        lfield<double, 9> xi_local_y(cnfg->Nxl, cnfg->Nyl);
        lfield<double, 9> A_local(cnfg->Nxl, cnfg->Nyl);
        lfield<double, 9> B_local(cnfg->Nxl, cnfg->Nyl);
        lfield<double, 9> uftmp(cnfg->Nxl, cnfg->Nyl);
        lfield<double, 9> uxiulocal_x(cnfg->Nxl, cnfg->Nyl);
        lfield<double, 9> uxiulocal_y(cnfg->Nxl, cnfg->Nyl);
        lfield<double, 9> uf_copy(cnfg->Nxl, cnfg->Nyl);
        lfield<double, 1> corr(cnfg->Nxl, cnfg->Nyl);
        gfield<double, 1> corr_global(Nx, Ny);
        std::vector<lfield<double, 1>*> sum, err;
        for (int m = 0; m < cnfg->measurements; m++) {
            sum.push_back(new lfield<double, 1>(cnfg->Nxl, cnfg->Nyl));
            err.push_back(new lfield<double, 1>(cnfg->Nxl, cnfg->Nyl));
        }
        const int every = cnfg->langevin_steps / cnfg->measurements > 0 ? cnfg->langevin_steps / cnfg->measurements : 1;

        // the reference measures after Langevin step l when l % every == 0, into
        // slot l * measurements / langevin_steps (integer division), several
        // measurements accumulating in one slot when measurements does not
        // divide langevin_steps (main_optimized.cpp:317-336).  every = 1 where
        // the reference would divide by zero (measurements > langevin_steps).
        auto measured = [&](int l) { return l % every == 0; };
        auto slot = [&](int l) { return (int)((long long)l * cnfg->measurements / cnfg->langevin_steps); };
        // steps from l up to and including the next measured one (or the end)
        auto chunk = [&](int l) {
            int n = 1;
            while (l + n - 1 < cnfg->langevin_steps - 1 && !measured(l + n - 1)) n++;
            return n;
        };
        auto measure = [&](lfield<double, 9>* U, int time) {
            uf_copy = *U;
            fourier2->execute2D(&uf_copy, 1);
            uf_copy.trace(&corr);
            corr_global.allgather(&corr, mpi);
            corr_global.average_and_symmetrize();
            corr_global.reduce(sum[time], err[time], mpi);
        };
        auto mv_init = [&](int stat) {
            jwhost::check(jw_noise_philox(jwhost::ctx(), cnfg->seed, (uint32_t)stat, 0), "noise stream");
            timespec starti;
            clock_gettime(CLOCK_MONOTONIC, &starti);
            uf.setToUnit();
            for (int i = 0; i < MVmodel->Ny_parameter; i++) {
                f.setMVModel(MVmodel);
                fourier2->execute2D(&f, 1);
                f.solvePoisson(cnfg->mass * pow(MVmodel->g_parameter, 2.0) * MVmodel->mu_parameter,
                               MVmodel->g_parameter, momtable);
                fourier2->execute2D(&f, 0);
                uf *= f;
            }
            jw_sync(jwhost::ctx());
            std::cout << "Initial condition time: " << seconds_since(starti) << std::endl;
        };

        if (cnfg->engine == ENGINE_BATCHED) {
            // all stat samples as one batch of trajectories (the library's batched form)
            timespec start;
            clock_gettime(CLOCK_MONOTONIC, &start);
            jw_batch* batch = nullptr;
            const int bN = Nx, bcount = cnfg->stat;
            jwhost::check(jw_batch_create(&batch, 0, 1, &bN, &bcount, cnfg->mixed ? 1 : 0), "batch");
            for (int stat = 0; stat < cnfg->stat; stat++) {
                printf("Gathering stat sample no. %i\n", stat);
                mv_init(stat);
                jwhost::check(jw_batch_set(batch, stat, uf.f), "batch set");
            }
            // each sample's stream continues after its MV sheets
            jwhost::check(jw_batch_philox(batch, cnfg->seed, 0, (uint32_t)MVmodel->Ny_parameter), "batch noise");
            for (int langevin = 0; langevin < cnfg->langevin_steps;) {
                const int n = chunk(langevin);
                timespec starte;
                clock_gettime(CLOCK_MONOTONIC, &starte);
                jwhost::check(jw_batch_evolve(batch, jwhost::coupling_code(cnfg->CouplingChoice),
                                              (int)cnfg->KernelChoice, cnfg->step, &jwhost::state().alpha, n,
                                              cnfg->reunit_every),
                              "batch evolve");
                jwhost::check(jw_batch_sync(batch), "batch evolve");
                std::cout << "Evolution time: " << seconds_since(starte) << " (" << n << " steps x " << cnfg->stat
                          << " trajectories)" << std::endl;
                langevin += n;
                if (measured(langevin - 1))
                    for (int stat = 0; stat < cnfg->stat; stat++) {
                        jwhost::check(jw_batch_get(batch, stat, uftmp.f), "batch get");
                        measure(&uftmp, slot(langevin - 1));
                    }
            }
            jw_batch_destroy(batch);
            std::cout << "Statistics time: " << seconds_since(start) << std::endl;
        } else {
            for (int stat = 0; stat < cnfg->stat; stat++) {
                timespec start;
                clock_gettime(CLOCK_MONOTONIC, &start);
                printf("Gathering stat sample no. %i\n", stat);
                mv_init(stat);
                uftmp = uf;
                for (int langevin = 0; langevin < cnfg->langevin_steps;) {
                    timespec starte;
                    clock_gettime(CLOCK_MONOTONIC, &starte);
                    int n = 1;
                    if (cnfg->engine == ENGINE_FUSED) {
                        n = chunk(langevin);
                        evolve(&uftmp, cnfg, n);
                    } else {
                        if (cnfg->CouplingChoice == NOISE_COUPLING_CONSTANT)
                            generate_gaussian_with_noise_coupling_constant(&xi_local_x, &xi_local_y, momtable, mpi,
                                                                           cnfg);
                        else
                            generate_gaussian(&xi_local_x, &xi_local_y, mpi, cnfg);
                        fourier2->execute2D(&xi_local_x, 1);
                        fourier2->execute2D(&xi_local_y, 1);
                        prepare_A_local(&A_local, &xi_local_x, &xi_local_y, momtable, mpi, cnfg->CouplingChoice,
                                        cnfg->KernelChoice);
                        fourier2->execute2D(&A_local, 0);
                        fourier2->execute2D(&xi_local_x, 0);
                        fourier2->execute2D(&xi_local_y, 0);
                        uxiulocal(&uxiulocal_x, &uxiulocal_y, &uftmp, &xi_local_x, &xi_local_y);
                        fourier2->execute2D(&uxiulocal_x, 1);
                        fourier2->execute2D(&uxiulocal_y, 1);
                        prepare_A_local(&B_local, &uxiulocal_x, &uxiulocal_y, momtable, mpi, cnfg->CouplingChoice,
                                        cnfg->KernelChoice);
                        fourier2->execute2D(&B_local, 0);
                        update_uf(&uftmp, &B_local, &A_local, cnfg->step);
                    }
                    jw_sync(jwhost::ctx());
                    std::cout << "Evolution time: " << seconds_since(starte) << std::endl;
                    langevin += n;
                    if (measured(langevin - 1)) measure(&uftmp, slot(langevin - 1));
                }
                std::cout << "Statistics time: " << seconds_since(start) << std::endl;
            }
        }
        for (int i = cnfg->measurements - 1; i < cnfg->measurements; i++)
            print(i, sum[i], err[i], momtable, 1.0 / 3.0 / cnfg->stat, mpi, cnfg->file_name);
        for (auto* p : sum) delete p;
        for (auto* p : err) delete p;
    }
    delete fourier2;
    delete MVmodel;
    delete momtable;
    delete mpi;
    delete cnfg;
    return 0;
}
