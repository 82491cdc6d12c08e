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
// batch = N:count,... runs size groups of uneven N together in one batch (the
// library's batched form): trajectory t (numbered over the groups) has the
// noise key {seed, t}, each group writes its own output file with its count as
// the stat normalisation.
//
// checkpoint_every = K (engines fused and reference) writes, every K Langevin
// steps, U and its noise state (<file_name>_ckpt.U, jw_checkpoint_write) and
// the driver's state (<file_name>_ckpt.state: stat sample, step, the sum/err
// accumulators); resume = 1 continues from them, bit for bit as the
// uninterrupted run; stop_after = S ends the run (without printing) after S
// Langevin steps, e.g. to fit a wall-clock allocation.
//
// Deviations from the reference, all stated: sum/err are sized by
// `measurements` (the reference sizes them by langevin_steps,
// main_optimized.cpp:134-135, 134 GB at N = 2048); the output file root is
// config file_name (the reference ignores it and writes "output_optimized");
// HATTA_COUPLING_CONSTANT is rejected; POSITION_EVOLUTION (N <= 256) correlates
// the noise of NOISE coupling with the per-mode square root of the reference's
// covariance instead of its dense Cholesky factor.
#include <stdio.h>
#include <string.h>
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

// the reference measures after Langevin step l when l % every == 0, into slot
// l * measurements / langevin_steps (integer division), several measurements
// accumulating in one slot when measurements does not divide langevin_steps
// (main_optimized.cpp:317-336).  every = 1 where the reference would divide
// by zero (measurements > langevin_steps).
struct Schedule {
    int steps, measurements, every, ckpt;
    explicit Schedule(const config* c)
        : steps(c->langevin_steps),
          measurements(c->measurements),
          every(c->langevin_steps / c->measurements > 0 ? c->langevin_steps / c->measurements : 1),
          ckpt(c->checkpoint_every) {}
    bool measured(int l) const { return l % every == 0; }
    int slot(int l) const { return (int)((long long)l * measurements / steps); }
    // steps from l up to and including the next measured one, stopping also
    // where a checkpoint falls (after steps that are multiples of ckpt), or the end
    int chunk(int l) const {
        int n = 1;
        while (l + n - 1 < steps - 1 && !measured(l + n - 1) && !(ckpt > 0 && (l + n) % ckpt == 0)) n++;
        return n;
    }
};

// driver state beside the U checkpoint: magic, stat sample, next Langevin
// step, measurements, V, then sum and err of every slot (2 V doubles each)
static const char kStateMagic[8] = {'J', 'W', 'D', 'R', 'V', '0', '0', '1'};

static void write_state(const std::string& path, int stat, int langevin, std::vector<lfield<double, 1>*>& sum,
                        std::vector<lfield<double, 1>*>& err) {
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) {
        fprintf(stderr, "checkpoint: cannot write %s\n", path.c_str());
        exit(3);
    }
    const int hdr[4] = {stat, langevin, (int)sum.size(), sum.empty() ? 0 : sum[0]->Nxl * sum[0]->Nyl};
    bool ok = fwrite(kStateMagic, 1, 8, f) == 8 && fwrite(hdr, sizeof hdr, 1, f) == 1;
    for (size_t m = 0; ok && m < sum.size(); m++)
        for (auto* fld : {sum[m], err[m]})
            ok = fwrite(fld->host_for_read(), sizeof(std::complex<double>), (size_t)hdr[3], f) == (size_t)hdr[3];
    if (fclose(f) != 0 || !ok) {
        fprintf(stderr, "checkpoint: writing %s failed\n", path.c_str());
        exit(3);
    }
}

static void read_state(const std::string& path, int* stat, int* langevin, std::vector<lfield<double, 1>*>& sum,
                       std::vector<lfield<double, 1>*>& err) {
    FILE* f = fopen(path.c_str(), "rb");
    char magic[8];
    int hdr[4];
    bool ok = f && fread(magic, 1, 8, f) == 8 && !memcmp(magic, kStateMagic, 8) && fread(hdr, sizeof hdr, 1, f) == 1 &&
              hdr[2] == (int)sum.size() && hdr[3] == sum[0]->Nxl * sum[0]->Nyl;
    for (size_t m = 0; ok && m < sum.size(); m++)
        for (auto* fld : {sum[m], err[m]})
            ok = fread(fld->host_for_write(), sizeof(std::complex<double>), (size_t)hdr[3], f) == (size_t)hdr[3];
    if (f) fclose(f);
    if (!ok) {
        fprintf(stderr, "resume: %s is missing or does not match this run\n", path.c_str());
        exit(3);
    }
    *stat = hdr[0];
    *langevin = hdr[1];
}

// batch = N:count,...: the size groups evolve together in one batch; one
// context per N for the MV starts and the measurements
static int run_batch_list(config* cnfg, mpi_class* mpi) {
    const int G = (int)cnfg->batch.size();
    std::vector<int> Ns, counts;
    for (auto& g : cnfg->batch) {
        Ns.push_back(g.first);
        counts.push_back(g.second);
    }
    const jw_alpha al{cnfg->lambda_L, cnfg->mu0_over_lambda, 0.2, 9.0};
    const Schedule sch(cnfg);
    std::vector<jw_ctx*> ctx(G, nullptr);
    std::vector<jw_field*> fld(G, nullptr);
    for (int g = 0; g < G; g++) {
        jwhost::check(jw_create(&ctx[g], Ns[g], 0), "group context");
        jwhost::check(jw_set_mv_zero_mode(ctx[g], cnfg->mv_zero_mode), "MV zero mode");
        jwhost::check(jw_field_create(ctx[g], 9, &fld[g]), "group field");
    }
    timespec start;
    clock_gettime(CLOCK_MONOTONIC, &start);
    jw_batch* batch = nullptr;
    jwhost::check(jw_batch_create(&batch, 0, G, Ns.data(), counts.data(), cnfg->mixed ? 1 : 0), "batch");
    jwhost::check(jw_batch_set_noise_space(batch, cnfg->noise_momentum ? JW_NOISE_MOMENTUM : JW_NOISE_POSITION),
                  "noise space");
    const int Ny = cnfg->elementaryWilsonLines;
    int t = 0;
    for (int g = 0; g < G; g++)
        for (int k = 0; k < counts[g]; k++, t++) {
            printf("Gathering trajectory no. %i (N = %i)\n", t, Ns[g]);
            // MV_class(1.0, mu/N, Ny) (main_optimized.cpp:88) and its sheets (178-191)
            jwhost::check(jw_noise_philox(ctx[g], cnfg->seed, (uint32_t)t, 0), "noise stream");
            jwhost::check(jw_mv_init(ctx[g], fld[g], Ny, 1.0, cnfg->mu / Ns[g], cnfg->mass, nullptr), "MV init");
            jwhost::check(jw_batch_set(batch, t, fld[g]), "batch set");
        }
    jwhost::check(jw_batch_philox(batch, cnfg->seed, 0, (uint32_t)Ny), "batch noise");
    // accumulators per group and slot (real parts; the values are real)
    std::vector<std::vector<std::vector<double>>> sum(G), err(G);
    for (int g = 0; g < G; g++) {
        const size_t V = (size_t)Ns[g] * Ns[g];
        sum[g].assign(cnfg->measurements, std::vector<double>(V, 0.0));
        err[g].assign(cnfg->measurements, std::vector<double>(V, 0.0));
    }
    std::vector<double> corr;
    for (int langevin = 0; langevin < cnfg->langevin_steps;) {
        const int n = sch.chunk(langevin);
        timespec starte;
        clock_gettime(CLOCK_MONOTONIC, &starte);
        jwhost::check(jw_batch_evolve(batch, jwhost::coupling_code(cnfg->CouplingChoice), (int)cnfg->KernelChoice,
                                      cnfg->step, &al, n, cnfg->reunit_every),
                      "batch evolve");
        jwhost::check(jw_batch_sync(batch), "batch evolve");
        std::cout << "Evolution time: " << seconds_since(starte) << " (" << n << " steps x " << t << " trajectories)"
                  << std::endl;
        langevin += n;
        if (!sch.measured(langevin - 1)) continue;
        const int s = sch.slot(langevin - 1);
        int tt = 0;
        for (int g = 0; g < G; g++) {
            corr.resize((size_t)Ns[g] * Ns[g]);
            for (int k = 0; k < counts[g]; k++, tt++) {
                jwhost::check(jw_batch_get(batch, tt, fld[g]), "batch get");
                jwhost::check(jw_measure_corr(ctx[g], fld[g], corr.data()), "measure");
                for (size_t i = 0; i < corr.size(); i++) {
                    sum[g][s][i] += corr[i];
                    err[g][s][i] += corr[i] * corr[i];
                }
            }
        }
    }
    jw_batch_destroy(batch);
    std::cout << "Statistics time: " << seconds_since(start) << std::endl;
    for (int g = 0; g < G; g++) {
        config cg = *cnfg;
        cg.Nx = cg.Ny = cg.Nxl = cg.Nyl = Ns[g];
        momenta mg(&cg, mpi);
        mg.set();
        const int m = cnfg->measurements - 1;
        print_rows(m, Ns[g], Ns[g], sum[g][m].data(), err[g][m].data(), &mg, 1.0 / 3.0 / counts[g], mpi,
                   cnfg->file_name);
        jw_field_destroy(fld[g]);
        jw_destroy(ctx[g]);
    }
    return 0;
}

int main(int argc, char* argv[]) {
    printf("INITIALIZATION\n");
    config* cnfg = new config;
    if (argc > 3 && !cnfg->read_config_from_file(argv[3])) return 1;
    if (cnfg->CouplingChoice == HATTA_COUPLING_CONSTANT) {
        fprintf(stderr, "HATTA_COUPLING_CONSTANT is not supported (the reference leaves its coupling unset)\n");
        return 2;
    }
    if (cnfg->EvolutionChoice == POSITION_EVOLUTION) {
        // the O(V^2) position-space step runs through the per-function calls
        if (cnfg->engine == ENGINE_BATCHED || cnfg->Nx > JW_POSITION_MAX_N) {
            fprintf(stderr, "POSITION_EVOLUTION runs the reference engine with N <= %d\n", JW_POSITION_MAX_N);
            return 2;
        }
        cnfg->engine = ENGINE_REFERENCE;
    }
    if (cnfg->engine == ENGINE_BATCHED && (cnfg->checkpoint_every > 0 || cnfg->resume || cnfg->stop_after > 0)) {
        fprintf(stderr, "checkpoints are written by the fused and reference engines\n");
        return 2;
    }
    mpi_class* mpi = new mpi_class(argc, argv);
    mpi->mpi_init(cnfg);
    mpi->mpi_exchange_grid();
    if (!cnfg->batch.empty()) {
        const int rc = run_batch_list(cnfg, mpi);
        delete mpi;
        delete cnfg;
        return rc;
    }
    const int Nx = cnfg->Nx, Ny = cnfg->Ny;

    momenta* momtable = new momenta(cnfg, mpi);
    momtable->set();
    positions postable(cnfg, mpi);
    postable.set();
    MV_class* MVmodel = new MV_class(1.0, cnfg->mu / Nx, cnfg->elementaryWilsonLines);
    fftw2D* fourier2 = new fftw2D(cnfg);
    fourier2->init2D();

    printf("ALLOCATION\n");
    int stopped = 0;
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
        const Schedule sch(cnfg);
        const std::string ckpt_U = cnfg->file_name + "_ckpt.U", ckpt_state = cnfg->file_name + "_ckpt.state";
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
            jwhost::check(jw_batch_set_noise_space(batch, cnfg->noise_momentum ? JW_NOISE_MOMENTUM : JW_NOISE_POSITION),
                          "noise space");
            for (int stat = 0; stat < cnfg->stat; stat++) {
                printf("Gathering stat sample no. %i\n", stat);
                mv_init(stat);
                jwhost::check(jw_batch_set(batch, stat, uf.dev_read()), "batch set");
            }
            // each sample's stream continues after its MV sheets
            jwhost::check(jw_batch_philox(batch, cnfg->seed, 0, (uint32_t)MVmodel->Ny_parameter), "batch noise");
            for (int langevin = 0; langevin < cnfg->langevin_steps;) {
                const int n = sch.chunk(langevin);
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
                if (sch.measured(langevin - 1))
                    for (int stat = 0; stat < cnfg->stat; stat++) {
                        jwhost::check(jw_batch_get(batch, stat, uftmp.dev_write()), "batch get");
                        measure(&uftmp, sch.slot(langevin - 1));
                    }
            }
            jw_batch_destroy(batch);
            std::cout << "Statistics time: " << seconds_since(start) << std::endl;
        } else {
            int stat0 = 0, langevin0 = 0, done = 0;
            if (cnfg->resume) {
                read_state(ckpt_state, &stat0, &langevin0, sum, err);
                printf("Resuming at stat sample %i, Langevin step %i\n", stat0, langevin0);
            }
            for (int stat = stat0; stat < cnfg->stat && !stopped; stat++) {
                timespec start;
                clock_gettime(CLOCK_MONOTONIC, &start);
                printf("Gathering stat sample no. %i\n", stat);
                int langevin = 0;
                if (cnfg->resume && stat == stat0 && langevin0 > 0) {
                    // U and its noise state at step langevin0 of this sample
                    jwhost::check(jw_checkpoint_read(jwhost::ctx(), uftmp.dev_write(), ckpt_U.c_str()), "resume");
                    langevin = langevin0;
                } else {
                    mv_init(stat);
                    uftmp = uf;
                }
                while (langevin < cnfg->langevin_steps) {
                    if (cnfg->stop_after > 0 && done >= cnfg->stop_after) {
                        stopped = 1;
                        break;
                    }
                    timespec starte;
                    clock_gettime(CLOCK_MONOTONIC, &starte);
                    int n = 1;
                    if (cnfg->engine == ENGINE_FUSED) {
                        n = sch.chunk(langevin);
                        evolve(&uftmp, cnfg, n);
                    } else if (cnfg->EvolutionChoice == POSITION_EVOLUTION) {
                        // main_optimized.cpp:243-247, 271-300: white noise, correlated
                        // for NOISE coupling, the position-space sums, the update
                        generate_gaussian(&xi_local_x, &xi_local_y, mpi, cnfg);
                        if (cnfg->CouplingChoice == NOISE_COUPLING_CONSTANT) {
                            multiplyByCholesky(&xi_local_x);
                            multiplyByCholesky(&xi_local_y);
                        }
                        prepare_A_and_B_local_all(&A_local, &B_local, &xi_local_x, &xi_local_y, &uftmp,
                                                  cnfg->CouplingChoice, cnfg->KernelChoice);
                        update_uf(&uftmp, &B_local, &A_local, cnfg->step);
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
                    done += n;
                    if (sch.measured(langevin - 1)) measure(&uftmp, sch.slot(langevin - 1));
                    if (cnfg->checkpoint_every > 0 && langevin % cnfg->checkpoint_every == 0 &&
                        langevin < cnfg->langevin_steps) {
                        jwhost::check(jw_checkpoint_write(jwhost::ctx(), uftmp.dev_read(), ckpt_U.c_str()),
                                      "checkpoint");
                        write_state(ckpt_state, stat, langevin, sum, err);
                    }
                }
                // a finished sample: the next one starts from its MV start
                if (!stopped && cnfg->checkpoint_every > 0 && stat + 1 < cnfg->stat)
                    write_state(ckpt_state, stat + 1, 0, sum, err);
                std::cout << "Statistics time: " << seconds_since(start) << std::endl;
            }
        }
        if (stopped)
            printf("Stopped after %i Langevin steps (stop_after); resume = 1 continues from the checkpoint\n",
                   cnfg->stop_after);
        else
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
