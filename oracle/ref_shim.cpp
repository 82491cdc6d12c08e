// oracle/ref_shim.cpp -- exposes the reference's own per-site code, compiled from
// its sources where they lie (su3_matrix.h, zheevh3-C-1.1/*.cpp, field.h,
// momenta.h; lattice fixed at 64 x 64 by the reference's config.h:31-32), as a
// C ABI in oracle/_ref/libjw_ref.so.
//
// TEST INFRASTRUCTURE ONLY: used by tests/ to pin the CPU oracle against the
// reference itself.  FFTW (needed by mpi_fftw_class.h and the drivers) is not in
// this image, so the transforms are not compiled here.
//
// All field arguments are 64 x 64 AoS arrays of complex<double> (field.h:199-209).

#include <cstring>
#include <new>
#include <complex>

#include "su3_matrix.h"
#include "config.h"
#include "field.h"

namespace {

// A single-rank mpi_class (pos_x = pos_y = rank = 0, size = 0) without calling
// MPI: the reference's null constructor (mpi_class.h:88-92) on zeroed storage.
// Built with -fno-lifetime-dse so the zeroing is kept.
mpi_class* single_rank() {
    alignas(mpi_class) static unsigned char storage[sizeof(mpi_class)];
    static mpi_class* m = nullptr;
    if (!m) {
        std::memset(storage, 0, sizeof storage);
        m = new (storage) mpi_class();
    }
    return m;
}

config* single_config() {
    static config* c = nullptr;
    if (!c) {
        c = new config;
        c->Nxl = Nx;
        c->Nyl = Ny;
    }
    return c;
}

momenta* single_momenta() {
    static momenta* mom = nullptr;
    if (!mom) {
        mom = new momenta(single_config(), single_rank());
        mom->set();
    }
    return mom;
}

template <int t>
void load(lfield<double, t>& f, const double* src) {
    std::memcpy((void*)f.u, src, sizeof(std::complex<double>) * t * Nx * Ny);
}

template <int t>
void store(const lfield<double, t>& f, double* dst) {
    std::memcpy(dst, (const void*)f.u, sizeof(std::complex<double>) * t * Nx * Ny);
}

}  // namespace

extern "C" {

int ref_lattice_size(void) { return Nx; }

// su3_matrix::exponentiate(sign), su3_matrix.h:173-253
int ref_su3_exponentiate(const double* in18, double sign, double* out18) {
    su3_matrix<double> a;
    for (int k = 0; k < 9; k++) a.m[k] = std::complex<double>(in18[2 * k], in18[2 * k + 1]);
    int st = a.exponentiate(sign);
    for (int k = 0; k < 9; k++) { out18[2 * k] = a.m[k].real(); out18[2 * k + 1] = a.m[k].imag(); }
    return st;
}

// function_zheevh3, su3_matrix.h:256-359
int ref_su3_eigen(const double* in18, double* Q18, double* w3) {
    std::complex<double> A[3][3], Q[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) A[i][j] = std::complex<double>(in18[2 * (3 * i + j)], in18[2 * (3 * i + j) + 1]);
    int st = function_zheevh3(A, Q, w3);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) { Q18[2 * (3 * i + j)] = Q[i][j].real(); Q18[2 * (3 * i + j) + 1] = Q[i][j].imag(); }
    return st;
}

// momenta::set, momenta.h:90-125
int ref_momenta(double* phat2, double* pbarx, double* pbary) {
    momenta* m = single_momenta();
    for (int i = 0; i < Nx * Ny; i++) {
        phat2[i] = m->phat2(i);
        pbarx[i] = m->pbarX(i);
        pbary[i] = m->pbarY(i);
    }
    return 0;
}

// prepare_A_local, field.h:1998-2075 (coupling: 0 SQRT, 1 NOISE; kernel: 0 LINEAR, 1 SIN)
int ref_prepare_A_local(double* A, const double* xix, const double* xiy, int coupling, int kernel) {
    if (coupling != 0 && coupling != 1) return -1;
    lfield<double, 9> a(Nx, Ny), x(Nx, Ny), y(Nx, Ny);
    load(x, xix);
    load(y, xiy);
    prepare_A_local(&a, &x, &y, single_momenta(), single_rank(),
                    coupling == 0 ? SQRT_COUPLING_CONSTANT : NOISE_COUPLING_CONSTANT,
                    kernel == 0 ? LINEAR_KERNEL : SIN_KERNEL);
    store(a, A);
    return 0;
}

// uxiulocal, field.h:1892-1933
int ref_uxiulocal(double* Vx, double* Vy, const double* U, const double* xix, const double* xiy) {
    lfield<double, 9> vx(Nx, Ny), vy(Nx, Ny), u(Nx, Ny), x(Nx, Ny), y(Nx, Ny);
    load(u, U);
    load(x, xix);
    load(y, xiy);
    uxiulocal(&vx, &vy, &u, &x, &y);
    store(vx, Vx);
    store(vy, Vy);
    return 0;
}

// update_uf, field.h:1966-1996
int ref_update_uf(double* U, const double* B, const double* A, double step) {
    lfield<double, 9> u(Nx, Ny), b(Nx, Ny), a(Nx, Ny);
    load(u, U);
    load(b, B);
    load(a, A);
    update_uf(&u, &b, &a, step);
    store(u, U);
    return 0;
}

// lfield::operator*= (U <- U exp(i f)), field.h:337-365
int ref_mul_exp(double* U, const double* F) {
    lfield<double, 9> u(Nx, Ny), f(Nx, Ny);
    load(u, U);
    load(f, F);
    u *= f;
    store(u, U);
    return 0;
}

// lfield::solvePoisson, field.h:867-877
int ref_solve_poisson(double* F, double mass, double g) {
    lfield<double, 9> f(Nx, Ny);
    load(f, F);
    f.solvePoisson(mass, g, single_momenta());
    store(f, F);
    return 0;
}

// lfield::setToUnit, field.h:196-211
int ref_set_unit(double* U) {
    lfield<double, 9> u(Nx, Ny);
    u.setToUnit();
    store(u, U);
    return 0;
}

// lfield::trace, field.h:1375-1432
int ref_trace(const double* U, double* corr) {
    lfield<double, 9> u(Nx, Ny);
    lfield<double, 1> c(Nx, Ny);
    load(u, U);
    u.trace(&c);
    store(c, corr);
    return 0;
}

// gfield::average_and_symmetrize, field.h:1460-1492
int ref_average_and_symmetrize(double* corr) {
    gfield<double, 1> g(Nx, Ny);
    std::memcpy((void*)g.u, corr, sizeof(std::complex<double>) * Nx * Ny);
    g.average_and_symmetrize();
    std::memcpy(corr, (const void*)g.u, sizeof(std::complex<double>) * Nx * Ny);
    return 0;
}

// lfield<double,1>::setCorrelationsForCouplingConstant(momenta*), field.h:1633-1651:
// alpha(phat2(k)) / sqrt(V) per mode, the input main_optimized.cpp:145-157
// transforms backward and factors densely (matrix.h:151-200)
int ref_correlations(double* corr) {
    lfield<double, 1> c(Nx, Ny);
    c.setCorrelationsForCouplingConstant(single_momenta());
    store(c, corr);
    return 0;
}

positions* single_positions() {
    static positions* pos = nullptr;
    if (!pos) {
        pos = new positions(single_config(), single_rank());
        pos->set();
    }
    return pos;
}

// prepare_A_and_B_local (field.h:2368-2519) for every site, the loop of
// main_optimized.cpp:288-297 (coupling: 0 SQRT, 1 NOISE; kernel: 0 LINEAR, 1 SIN)
int ref_prepare_A_and_B_local(double* A, double* B, const double* xix, const double* xiy, const double* U,
                              int coupling, int kernel) {
    if (coupling != 0 && coupling != 1) return -1;
    gfield<double, 9> x(Nx, Ny), y(Nx, Ny), u(Nx, Ny);
    lfield<double, 9> a(Nx, Ny), b(Nx, Ny);
    std::memcpy((void*)x.u, xix, sizeof(std::complex<double>) * 9 * Nx * Ny);
    std::memcpy((void*)y.u, xiy, sizeof(std::complex<double>) * 9 * Nx * Ny);
    std::memcpy((void*)u.u, U, sizeof(std::complex<double>) * 9 * Nx * Ny);
    a.setToZero();
    b.setToZero();
    for (int xx = 0; xx < Nx; xx++)
        for (int yy = 0; yy < Ny; yy++)
            prepare_A_and_B_local(xx, yy, xx, yy, &x, &y, &a, &b, &u, single_positions(), 0,
                                  coupling == 0 ? SQRT_COUPLING_CONSTANT : NOISE_COUPLING_CONSTANT,
                                  kernel == 0 ? LINEAR_KERNEL : SIN_KERNEL);
    store(a, A);
    store(b, B);
    return 0;
}

// gfield::reduce(sum, err, mpi), field.h:1521-1536
int ref_reduce(const double* corr, double* sum, double* err) {
    gfield<double, 1> g(Nx, Ny);
    lfield<double, 1> s(Nx, Ny), e(Nx, Ny);
    std::memcpy((void*)g.u, corr, sizeof(std::complex<double>) * Nx * Ny);
    load(s, sum);
    load(e, err);
    g.reduce(&s, &e, single_rank());
    store(s, sum);
    store(e, err);
    return 0;
}

}  // extern "C"
