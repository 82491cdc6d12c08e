// field.h -- lfield / gfield and the free functions of the Langevin step, the
// surface of the reference's field.h (lfield/gfield field.h:101-299, the step
// functions field.h:1892-2075 and 2116-2365, measurement field.h:535-591 and
// 1375-1536, print field.h:2521-2547), over the HIP library's C ABI.
//
// A field lives in device memory (jw_field).  `u` is its host view, kept
// coherent with the device copy: reading u[i] after a device operation on the
// field downloads the field first, writing u[i] marks the host copy as newer,
// and the next device operation that reads the field uploads it first (the
// reference's `u` is the live host storage, field.h:70-80).  A raw pointer
// taken from u (std::complex<T>* p = field.u) stays valid until the next
// device operation on the field.  All fields of a run share one context (one
// lattice on one GPU), created by fftw2D::init2D().
// Errors print the library's message and exit with status 3: the reference
// prints and carries on (e.g. field.h:749-753), which hides failures.
#ifndef JW_HOST_FIELD_H
#define JW_HOST_FIELD_H

#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include <complex>
#include <string>
#include <vector>

#include "jimwlk.h"
#include "config.h"
#include "lattice.h"

namespace jwhost {

struct State {
    jw_ctx* ctx = nullptr;
    config* cnfg = nullptr;
    jw_alpha alpha{6.0, 2.5, 0.2, 9.0};
};

inline State& state() {
    static State s;
    return s;
}

inline void check(int rc, const char* what) {
    if (rc != JW_OK) {
        fprintf(stderr, "%s failed (%d): %s\n", what, rc, jw_last_error());
        exit(3);
    }
}

inline jw_ctx* ctx() {
    if (!state().ctx) {
        fprintf(stderr, "fields need the lattice context: construct fftw2D and call init2D() first\n");
        exit(3);
    }
    return state().ctx;
}

inline int coupling_code(Coupling c) { return (int)c; }

}  // namespace jwhost

template <class T, int t>
class lfield;

// the host view `u` of an lfield: element access and conversion to a raw
// pointer through the owner's coherence calls
template <class T, int t>
class host_view {
    lfield<T, t>* owner;

   public:
    explicit host_view(lfield<T, t>* o) : owner(o) {}
    host_view(const host_view&) = delete;
    host_view& operator=(const host_view&) = delete;
    std::complex<T>& operator[](size_t i) { return owner->host_for_write()[i]; }
    const std::complex<T>& operator[](size_t i) const { return owner->host_for_read()[i]; }
    operator std::complex<T>*() { return owner->host_for_write(); }
    operator const std::complex<T>*() const { return owner->host_for_read(); }
};

template <class T, int t>
class lfield {
    jw_field* f = nullptr;
    std::vector<std::complex<T>> h;   // host copy behind `u`
    bool host_valid = false;          // h equals the device copy
    bool host_newer = false;          // h was written since the last upload

   public:
    int Nxl, Nyl;
    host_view<T, t> u;

    lfield(int NNx, int NNy) : Nxl(NNx), Nyl(NNy), u(this) {
        static_assert(t == 1 || t == 9, "fields have 1 or 9 components");
        jwhost::check(jw_field_create(jwhost::ctx(), t, &f), "lfield allocation");
    }
    lfield(const lfield& in) : Nxl(in.Nxl), Nyl(in.Nyl), u(this) {
        jwhost::check(jw_field_create(jwhost::ctx(), t, &f), "lfield allocation");
        jwhost::check(jw_field_copy(f, in.dev_read()), "lfield copy");
    }
    virtual ~lfield() { jw_field_destroy(f); }

    // the device field for an operation that reads it (host writes uploaded first)
    jw_field* dev_read() const {
        lfield* self = const_cast<lfield*>(this);
        if (self->host_newer) {
            jwhost::check(jw_field_upload(f, reinterpret_cast<const double*>(h.data())), "upload");
            self->host_newer = false;
        }
        return f;
    }
    // the device field for an operation that writes it (the host copy goes stale)
    jw_field* dev_write() {
        dev_read();
        host_valid = false;
        return f;
    }
    // the host copy, current with the device
    const std::complex<T>* host_for_read() const {
        lfield* self = const_cast<lfield*>(this);
        if (!self->host_valid) {
            self->h.resize((size_t)t * Nxl * Nyl);
            jwhost::check(jw_field_download(f, reinterpret_cast<double*>(self->h.data())), "download");
            self->host_valid = true;
        }
        return h.data();
    }
    std::complex<T>* host_for_write() {
        host_for_read();
        host_newer = true;
        return h.data();
    }
    // explicit synchronisation: download now / upload now
    std::complex<T>* to_host() {
        host_valid = false;
        return host_for_write();
    }
    int from_host() {
        if (host_valid) host_newer = true;
        dev_read();
        return 1;
    }

    int loc_pos(int x, int y) { return y + Nyl * x; }

    int setToZero() {
        jwhost::check(jw_set_to_zero(dev_write()), "setToZero");
        return 1;
    }
    int setToUnit() {
        jwhost::check(jw_set_to_unit(dev_write()), "setToUnit");
        return 1;
    }
    lfield& operator=(const lfield& in) {
        if (this != &in) jwhost::check(jw_field_copy(dev_write(), in.dev_read()), "operator=");
        return *this;
    }
    // U <- U exp(i f)   (field.h:337-365)
    lfield& operator*=(const lfield& g) {
        if (this != &g) jwhost::check(jw_mul_exp(jwhost::ctx(), dev_write(), g.dev_read()), "operator*=");
        return *this;
    }
    // one MV sheet of colour charge (field.h:683-757), charges from the context's noise source
    int setMVModel(MV_class* mv) {
        jwhost::check(jw_set_mv_model(jwhost::ctx(), dev_write(), mv->g_parameter, mv->mu_parameter, mv->Ny_parameter),
                      "setMVModel");
        return 1;
    }
    int solvePoisson(double mass, double g, momenta*) {
        jwhost::check(jw_solve_poisson(jwhost::ctx(), dev_write(), mass, g), "solvePoisson");
        return 1;
    }
    int trace(lfield<double, 1>* cc) {
        jwhost::check(jw_trace(jwhost::ctx(), dev_read(), cc->dev_write()), "trace");
        return 1;
    }
};

template <class T, int t>
class gfield : public lfield<T, t> {
   public:
    int Nxg, Nyg;
    gfield(int NNx, int NNy) : lfield<T, t>(NNx, NNy), Nxg(NNx), Nyg(NNy) {}
    // one rank holds the whole lattice: the gather is a device copy (field.h:535-591)
    int allgather(lfield<T, t>* ulocal, mpi_class*) {
        jwhost::check(jw_field_copy(this->dev_write(), ulocal->dev_read()), "allgather");
        return 1;
    }
    int average_and_symmetrize() {
        jwhost::check(jw_average_and_symmetrize(jwhost::ctx(), this->dev_write()), "average_and_symmetrize");
        return 1;
    }
    int reduce(lfield<T, t>* sum, lfield<T, t>* err, mpi_class*) {
        jwhost::check(jw_reduce(jwhost::ctx(), this->dev_read(), sum->dev_write(), err->dev_write()), "reduce");
        return 1;
    }
};

// fftw2D (mpi_fftw_class.h:113-352): unitary 2-D transforms of fields.  init2D
// also creates the lattice context on device 0 and seeds the Philox stream.
class fftw2D {
    config* cnfg;

   public:
    explicit fftw2D(config* c) : cnfg(c) {}
    ~fftw2D() {
        jwhost::State& s = jwhost::state();
        if (s.ctx) jw_destroy(s.ctx);
        s.ctx = nullptr;
    }
    int init2D() {
        jwhost::State& s = jwhost::state();
        if (cnfg->Nx != cnfg->Ny) {
            fprintf(stderr, "fftw2D: Nx == Ny required (%d x %d)\n", cnfg->Nx, cnfg->Ny);
            exit(3);
        }
        if (!s.ctx) jwhost::check(jw_create(&s.ctx, cnfg->Nx, 0), "lattice context");
        s.cnfg = cnfg;
        s.alpha = jw_alpha{cnfg->lambda_L, cnfg->mu0_over_lambda, 0.2, 9.0};
        jwhost::check(jw_noise_philox(s.ctx, cnfg->seed, 0, 0), "noise stream");
        jwhost::check(jw_set_mv_zero_mode(s.ctx, cnfg->mv_zero_mode), "MV zero mode");
        jwhost::check(jw_set_noise_space(s.ctx, cnfg->noise_momentum ? JW_NOISE_MOMENTUM : JW_NOISE_POSITION),
                      "noise space");
        return 1;
    }
    template <class T, int t>
    int execute2D(lfield<T, t>* f, int dir) {
        jwhost::check(jw_execute2D(jwhost::ctx(), f->dev_write(), dir), "execute2D");
        return 1;
    }
};

// ---------------------------------------------------------------------------
// the step functions (field.h:1892-2075, 2116-2365)
// ---------------------------------------------------------------------------
template <class T, int t>
int generate_gaussian(lfield<T, t>* xi_x, lfield<T, t>* xi_y, mpi_class*, config*) {
    jwhost::check(jw_generate_gaussian(jwhost::ctx(), xi_x->dev_write(), xi_y->dev_write()), "generate_gaussian");
    return 1;
}

template <class T, int t>
int generate_gaussian_with_noise_coupling_constant(lfield<T, t>* xi_x, lfield<T, t>* xi_y, momenta*, mpi_class*,
                                                    config*) {
    jwhost::check(jw_generate_gaussian_with_noise_coupling_constant(jwhost::ctx(), xi_x->dev_write(), xi_y->dev_write(),
                                                                    &jwhost::state().alpha),
                  "generate_gaussian_with_noise_coupling_constant");
    return 1;
}

template <class T, int t>
int prepare_A_local(lfield<T, t>* A, lfield<T, t>* xi_x, lfield<T, t>* xi_y, momenta*, mpi_class*, Coupling p,
                    Kernel kk) {
    jwhost::check(jw_prepare_A_local(jwhost::ctx(), A->dev_write(), xi_x->dev_read(), xi_y->dev_read(),
                                     jwhost::coupling_code(p), (int)kk,
                                     &jwhost::state().alpha),
                  "prepare_A_local");
    return 1;
}

template <class T, int t>
int uxiulocal(lfield<T, t>* vx, lfield<T, t>* vy, lfield<T, t>* uf, lfield<T, t>* xi_x, lfield<T, t>* xi_y) {
    jwhost::check(jw_uxiulocal(jwhost::ctx(), vx->dev_write(), vy->dev_write(), uf->dev_read(), xi_x->dev_read(),
                               xi_y->dev_read()),
                  "uxiulocal");
    return 1;
}

template <class T, int t>
int update_uf(lfield<T, t>* uf, lfield<T, t>* B, lfield<T, t>* A, double step) {
    jwhost::check(jw_update_uf(jwhost::ctx(), uf->dev_write(), B->dev_read(), A->dev_read(), step), "update_uf");
    return 1;
}

// POSITION_EVOLUTION (main_optimized.cpp:271-297): the reference calls
// prepare_A_and_B_local (field.h:2368-2519) once per site over gathered global
// fields; here one call computes every site (the gather is the identity on one
// GPU).  N <= JW_POSITION_MAX_N.
template <class T, int t>
int prepare_A_and_B_local_all(lfield<T, t>* A, lfield<T, t>* B, lfield<T, t>* xi_x, lfield<T, t>* xi_y,
                              lfield<T, t>* uf, Coupling p, Kernel kk) {
    jwhost::check(jw_prepare_A_and_B_local(jwhost::ctx(), A->dev_write(), B->dev_write(), xi_x->dev_read(),
                                           xi_y->dev_read(), uf->dev_read(), jwhost::coupling_code(p), (int)kk),
                  "prepare_A_and_B_local");
    return 1;
}

// gfield::multiplyByCholesky (field.h:1653-1689) with the coupling covariance
// of main_optimized.cpp:145-157: the per-mode square root of the same
// covariance (jw_correlate_noise)
template <class T, int t>
int multiplyByCholesky(lfield<T, t>* xi) {
    jwhost::check(jw_correlate_noise(jwhost::ctx(), xi->dev_write(), &jwhost::state().alpha), "multiplyByCholesky");
    return 1;
}

// nsteps of the momentum-space step in the library's generator-basis engine
// (same result as the per-function sequence above, to rounding), in FP64 or in
// the mixed-precision form (config precision = mixed).
template <class T, int t>
int evolve(lfield<T, t>* uf, config* cnfg, int nsteps) {
    const int c = jwhost::coupling_code(cnfg->CouplingChoice);
    if (cnfg->mixed)
        jwhost::check(jw_evolve_mixed(jwhost::ctx(), uf->dev_write(), c, (int)cnfg->KernelChoice, cnfg->step,
                                      &jwhost::state().alpha, nsteps, cnfg->reunit_every),
                      "evolve (mixed precision)");
    else
        jwhost::check(jw_evolve(jwhost::ctx(), uf->dev_write(), c, (int)cnfg->KernelChoice, cnfg->step, &jwhost::state().alpha,
                                nsteps),
                      "evolve");
    return 1;
}

// print (field.h:2521-2547): rows "meas kx ky sqrt(phat2) x*phat2*sum x*phat2*err"
// for |kx - ky| <= 4, appended to <fileroot>_<Nx>_<Ny>_mpi<size>_r<rank>.dat.
// sum and err: the real parts of the accumulated t = 1 fields, Nxl x Nyl.
inline int print_rows(int measurement, int Nxl, int Nyl, const double* sum, const double* err, momenta* mom, double x,
                      mpi_class* mpi, const std::string& fileroot) {
    char filename[512];
    snprintf(filename, sizeof filename, "%s_%i_%i_mpi%i_r%i.dat", fileroot.c_str(), Nxl, Nyl, mpi->getSize(),
             mpi->getRank());
    FILE* f = fopen(filename, "a+");
    if (!f) {
        fprintf(stderr, "print: cannot open %s\n", filename);
        return 0;
    }
    for (int xx = 0; xx < Nxl; xx++)
        for (int yy = 0; yy < Nyl; yy++) {
            const int i = xx * Nyl + yy;
            if (fabs((double)(xx - yy)) <= 4)
                fprintf(f, "%i %i %i \t %f %e %e\n", measurement, xx, yy, sqrt(mom->phat2(i)),
                        x * (mom->phat2(i)) * sum[i], x * (mom->phat2(i)) * err[i]);
        }
    fclose(f);
    return 1;
}

template <class T, int t>
int print(int measurement, lfield<T, t>* sum, lfield<T, t>* err, momenta* mom, double x, mpi_class* mpi,
          const std::string& fileroot) {
    const std::complex<T>* s = sum->host_for_read();
    const std::complex<T>* e = err->host_for_read();
    std::vector<double> sr((size_t)sum->Nxl * sum->Nyl), er(sr.size());
    for (size_t i = 0; i < sr.size(); i++) {
        sr[i] = s[i * t].real();
        er[i] = e[i * t].real();
    }
    return print_rows(measurement, sum->Nxl, sum->Nyl, sr.data(), er.data(), mom, x, mpi, fileroot);
}

#endif
