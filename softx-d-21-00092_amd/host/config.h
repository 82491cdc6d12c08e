// config.h -- run parameters and their file parser, the surface of the
// reference's config.h:28-139 with the lattice size made a run-time value.
//
// The file format is the reference's: eleven "key = value" lines in a fixed
// order (config.h:76-136; example file `config`).  As in the reference, mu,
// mass and step are read through a float (config.h:79-93), so 1e-4 becomes
// 9.99999974737875e-05.  Optional lines after the eleven select what the
// reference fixes at compile time or does not have:
//   N = 2048                lattice extent (reference: #define Nx/Ny 64)
//   seed = 1                Philox key (the reference seeds from the clock)
//   lambda_L = 6            Lambda*N of the running coupling (field.h:2273)
//   mu0_over_lambda = 2.5   mu0/Lambda (field.h:2257)
//   engine = fused          fused | reference (the per-function calls) |
//                           batched (all stat samples evolve together)
//   precision = fp64        fp64 | mixed (FP32 fields and FFTs, FP64 exponentials)
//   reunit_every = 10       mixed precision: steps between projections onto SU(3)
//   float_parse = 1         1: mu, mass and step pass through a float as in the
//                           reference; 0: they are read as doubles
//   mv_zero_mode = 0        0: the reference's k = 0 factor -g/m^2 in solvePoisson;
//                           1: each MV sheet's k = 0 mode is set to zero
//   noise_space = position  position: 16 Philox normals per site (the reference's
//                           generate_gaussian); momentum: the noise spectra drawn
//                           per mode (same distribution, no noise pass; fused
//                           and batched engines, couplings other than NOISE)
//   batch = 256:48,512:32   batched engine: size groups N:count (uneven sizes);
//                           each group writes its own output file
//   checkpoint_every = 0    steps between checkpoints of (U, noise state) per
//                           stat sample (0: none); files <file_name>_ckpt_<stat>
//   resume = 0              1: continue from the checkpoint files
//   stop_after = 0          end the run after this many Langevin steps (0: never),
//                           leaving the checkpoint for resume
#ifndef JW_HOST_CONFIG_H
#define JW_HOST_CONFIG_H

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <utility>
#include <vector>

enum Evolution { POSITION_EVOLUTION, MOMENTUM_EVOLUTION };

// The reference's three choices (config.h:40) and two additions: FIXED_COUPLING
// (kernel factor 1, white noise) and CORRELATED_NOISE_COUPLING (per-mode
// sqrt(alpha(k)) on the noise spectrum).
enum Coupling {
    SQRT_COUPLING_CONSTANT,
    NOISE_COUPLING_CONSTANT,
    HATTA_COUPLING_CONSTANT,
    FIXED_COUPLING,
    CORRELATED_NOISE_COUPLING
};

enum Kernel { LINEAR_KERNEL, SIN_KERNEL };

enum Engine { ENGINE_FUSED, ENGINE_REFERENCE, ENGINE_BATCHED };

class config {
   public:
    int Nx = 64, Ny = 64;
    int Nxl = 0, Nyl = 0;
    int Nxl_buf = 0, Nyl_buf = 0;
    int proc_x = 1, proc_y = 1;

    unsigned seed = 1;
    int stat = 4;
    double mu = 30.72;
    double mass = 0.0001;
    int elementaryWilsonLines = 50;
    std::string file_name = "output";
    int langevin_steps = 100;
    int measurements = 100;
    double step = 0.0004;

    Evolution EvolutionChoice = MOMENTUM_EVOLUTION;
    Coupling CouplingChoice = NOISE_COUPLING_CONSTANT;
    Kernel KernelChoice = SIN_KERNEL;

    double lambda_L = 6.0;
    double mu0_over_lambda = 2.5;
    Engine engine = ENGINE_FUSED;
    bool mixed = false;
    int reunit_every = 10;
    bool float_parse = true;
    int mv_zero_mode = 0;
    bool noise_momentum = false;
    std::vector<std::pair<int, int>> batch;     // (N, count) size groups
    int checkpoint_every = 0;
    bool resume = false;
    int stop_after = 0;

    // Returns 1 on success, 0 when the file cannot be read or a required key is
    // missing (the reference returns nothing and reads garbage).
    int read_config_from_file(const std::string& path) {
        FILE* f = fopen(path.c_str(), "r");
        if (!f) {
            fprintf(stderr, "config: cannot open %s\n", path.c_str());
            return 0;
        }
        // mu, mass and step are kept as text until float_parse is known
        char buf[256], tmu[128], tmass[128], tstep[128];
        int ok = 1;
        ok &= fscanf(f, " stat = %i", &stat) == 1;
        ok &= fscanf(f, " mu = %127s", tmu) == 1;
        ok &= fscanf(f, " mass = %127s", tmass) == 1;
        ok &= fscanf(f, " elementaryWilsonLines = %i", &elementaryWilsonLines) == 1;
        ok &= fscanf(f, " file_name = %255s", buf) == 1;
        file_name = buf;
        ok &= fscanf(f, " langevin_steps = %i", &langevin_steps) == 1;
        ok &= fscanf(f, " measurements = %i", &measurements) == 1;
        ok &= fscanf(f, " step = %127s", tstep) == 1;
        char evolution[100], coupling[100], kernel[100];
        ok &= fscanf(f, " EvolutionChoice = %99s", evolution) == 1;
        ok &= fscanf(f, " CouplingChoice = %99s", coupling) == 1;
        ok &= fscanf(f, " KernelChoice = %99s", kernel) == 1;
        if (!ok) {
            fprintf(stderr, "config: %s does not hold the eleven reference keys in order\n", path.c_str());
            fclose(f);
            return 0;
        }
        if (!strcmp(evolution, "MOMENTUM_EVOLUTION")) EvolutionChoice = MOMENTUM_EVOLUTION;
        if (!strcmp(evolution, "POSITION_EVOLUTION")) EvolutionChoice = POSITION_EVOLUTION;
        if (!strcmp(coupling, "SQRT_COUPLING_CONSTANT")) CouplingChoice = SQRT_COUPLING_CONSTANT;
        if (!strcmp(coupling, "NOISE_COUPLING_CONSTANT")) CouplingChoice = NOISE_COUPLING_CONSTANT;
        if (!strcmp(coupling, "HATTA_COUPLING_CONSTANT")) CouplingChoice = HATTA_COUPLING_CONSTANT;
        if (!strcmp(coupling, "FIXED_COUPLING")) CouplingChoice = FIXED_COUPLING;
        if (!strcmp(coupling, "CORRELATED_NOISE_COUPLING")) CouplingChoice = CORRELATED_NOISE_COUPLING;
        if (!strcmp(kernel, "LINEAR_KERNEL")) KernelChoice = LINEAR_KERNEL;
        if (!strcmp(kernel, "SIN_KERNEL")) KernelChoice = SIN_KERNEL;
        // optional keys, any order
        char key[64], val[128];
        while (fscanf(f, " %63[^= ] = %127s", key, val) == 2) {
            if (!strcmp(key, "N")) Nx = Ny = atoi(val);
            else if (!strcmp(key, "seed")) seed = (unsigned)strtoul(val, nullptr, 10);
            else if (!strcmp(key, "lambda_L")) lambda_L = atof(val);
            else if (!strcmp(key, "mu0_over_lambda")) mu0_over_lambda = atof(val);
            else if (!strcmp(key, "engine")) {
                if (!strcmp(val, "reference")) engine = ENGINE_REFERENCE;
                else if (!strcmp(val, "batched")) engine = ENGINE_BATCHED;
                else if (!strcmp(val, "fused")) engine = ENGINE_FUSED;
                else {
                    fprintf(stderr, "config: engine %s: expected fused, reference or batched\n", val);
                    ok = 0;
                }
            } else if (!strcmp(key, "precision")) {
                if (!strcmp(val, "mixed")) mixed = true;
                else if (!strcmp(val, "fp64")) mixed = false;
                else {
                    fprintf(stderr, "config: precision %s: expected fp64 or mixed\n", val);
                    ok = 0;
                }
            } else if (!strcmp(key, "reunit_every")) reunit_every = atoi(val);
            else if (!strcmp(key, "float_parse")) float_parse = atoi(val) != 0;
            else if (!strcmp(key, "mv_zero_mode")) mv_zero_mode = atoi(val);
            else if (!strcmp(key, "noise_space")) {
                if (!strcmp(val, "momentum")) noise_momentum = true;
                else if (!strcmp(val, "position")) noise_momentum = false;
                else {
                    fprintf(stderr, "config: noise_space %s: expected position or momentum\n", val);
                    ok = 0;
                }
            }
            else if (!strcmp(key, "checkpoint_every")) checkpoint_every = atoi(val);
            else if (!strcmp(key, "resume")) resume = atoi(val) != 0;
            else if (!strcmp(key, "stop_after")) stop_after = atoi(val);
            else if (!strcmp(key, "batch")) ok &= parse_batch(val);
            else fprintf(stderr, "config: unknown key %s ignored\n", key);
        }
        fclose(f);
        // the reference's fscanf("%f") into a float (config.h:79-93), or doubles
        ok &= to_real(tmu, &mu) && to_real(tmass, &mass) && to_real(tstep, &step);
        if (!batch.empty()) engine = ENGINE_BATCHED;
        if (checkpoint_every < 0 || stop_after < 0 || (mv_zero_mode != 0 && mv_zero_mode != 1)) {
            fprintf(stderr, "config: checkpoint_every and stop_after must be >= 0, mv_zero_mode 0 or 1\n");
            ok = 0;
        }
        if (engine == ENGINE_REFERENCE && mixed) {
            fprintf(stderr, "config: the reference engine runs in FP64 only\n");
            ok = 0;
        }
        if (engine == ENGINE_REFERENCE && noise_momentum) {
            fprintf(stderr, "config: the reference engine draws position-space noise\n");
            ok = 0;
        }
        return ok;
    }

   private:
    int to_real(const char* t, double* out) const {
        char* end = nullptr;
        if (float_parse) *out = (double)strtof(t, &end);
        else *out = strtod(t, &end);
        if (end == t || *end) {
            fprintf(stderr, "config: %s is not a number\n", t);
            return 0;
        }
        return 1;
    }

    // "N:count,N:count,..."
    int parse_batch(const char* v) {
        batch.clear();
        const char* p = v;
        while (*p) {
            int n = 0, cnt = 0, used = 0;
            if (sscanf(p, "%d:%d%n", &n, &cnt, &used) != 2 || n < 1 || cnt < 1) {
                fprintf(stderr, "config: batch %s: expected N:count[,N:count...]\n", v);
                return 0;
            }
            batch.emplace_back(n, cnt);
            p += used;
            if (*p == ',') p++;
            else if (*p) {
                fprintf(stderr, "config: batch %s: expected N:count[,N:count...]\n", v);
                return 0;
            }
        }
        return !batch.empty();
    }
};

#endif
