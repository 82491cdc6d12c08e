// oracle/jw_oracle.cpp -- CPU restatement of the reference JIMWLK momentum-space
// Langevin path (main_optimized.cpp:159-355 of ElsevierSoftwareX/SOFTX-D-21-00092).
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle the HIP library is checked
// against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
// may load it; the product path (the HIP library and the host classes above it)
// never calls it.
//
// Fields are host arrays in the reference's AoS layout (field.h:70-80,
// 199-209): complex<double> u[(x*N + y)*t + k], k = 3*row + col, stored as
// interleaved re/im doubles.  The lattice is N x N with N a power of two.
// Every function returns 0 on success and a negative value on bad arguments.
//
// Noise is injected: the caller passes the standard normals the reference would
// have drawn, in the reference's draw order (field.h:2162-2165: n0,m0,n1,m1,...;
// field.h:708-710 for the MV charges: n0..n7).  The Philox stream used by the HIP
// library is restated here too (Salmon et al., SC'11), so the two can be compared.
//
// Each function below names the reference lines it restates; the tests that pin
// it are in tests/test_oracle.py.

#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <cfloat>
#include <vector>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef std::complex<double> cd;

namespace {

inline cd* C(double* p) { return reinterpret_cast<cd*>(p); }
inline const cd* C(const double* p) { return reinterpret_cast<const cd*>(p); }

bool pow2(int n) { return n >= 2 && (n & (n - 1)) == 0; }

// ---------------------------------------------------------------------------
// 3x3 complex algebra, row-major.  Products accumulate k = 0,1,2 starting from
// zero, as su3_matrix::operator* does (su3_matrix.h:155-168).
// ---------------------------------------------------------------------------
struct M3 { cd m[9]; };

inline M3 mul(const M3& a, const M3& b) {
    M3 c;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            cd s = 0.0;
            for (int k = 0; k < 3; k++) s += a.m[3 * i + k] * b.m[3 * k + j];
            c.m[3 * i + j] = s;
        }
    return c;
}

inline M3 dagger(const M3& a) {
    M3 c;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) c.m[3 * i + j] = std::conj(a.m[3 * j + i]);
    return c;
}

inline double abs2(cd z) { return z.real() * z.real() + z.imag() * z.imag(); }

// ---------------------------------------------------------------------------
// Hermitian 3x3 eigen-decomposition, Kopp's hybrid method (J. Kopp, "Efficient
// numerical diagonalization of hermitian 3x3 matrices", Int. J. Mod. Phys. C 19
// (2008) 523, arXiv:physics/0610206), as the reference uses it through
// function_zheevh3 (su3_matrix.h:256-359) and zheevc3 (zheevc3.cpp:33-86).
// Only the upper triangle and the real part of the diagonal are read
// (su3_matrix.h:263-264).
// ---------------------------------------------------------------------------

// Cardano's formula for the roots of det(A - w) = 0 (paper, Sec. 2.1).
void cardano_eigenvalues(const cd A[3][3], double w[3]) {
    const double a = A[0][0].real(), b = A[1][1].real(), c = A[2][2].real();
    const cd d = A[0][1], e = A[1][2], f = A[0][2];
    const cd de = d * e;
    const double dd = abs2(d), ee = abs2(e), ff = abs2(f);
    // characteristic polynomial  w^3 - m w^2 + c1 w + c0
    const double m = a + b + c;
    const double c1 = (a * b + a * c + b * c) - (dd + ee + ff);
    const double c0 = c * dd + a * ee + b * ff - a * b * c
                      - 2.0 * (f.real() * de.real() + f.imag() * de.imag());
    const double p = m * m - 3.0 * c1;
    const double q = m * (p - 1.5 * c1) - 13.5 * c0;
    const double sp = std::sqrt(std::fabs(p));
    double disc = 27.0 * (0.25 * c1 * c1 * (p - c1) + c0 * (q + 6.75 * c0));
    const double phi = (1.0 / 3.0) * std::atan2(std::sqrt(std::fabs(disc)), q);
    const double cs = sp * std::cos(phi);
    const double sn = (1.0 / 1.73205080756887729352744634151) * sp * std::sin(phi);
    const double mid = (1.0 / 3.0) * (m - cs);
    w[0] = mid + cs;
    w[2] = mid + sn;
    w[1] = mid - sn;
}

// Cyclic complex Jacobi sweeps.  Used where the cross-product eigenvectors
// are unreliable (the reference falls back to Householder + QL there,
// zheevq3.cpp:31-137; any accurate solver gives the same exp(iH) to rounding).
void jacobi_eigen(const cd A[3][3], cd Q[3][3], double w[3]) {
    cd H[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            if (i == j) H[i][j] = A[i][i].real();
            else if (i < j) H[i][j] = A[i][j];
            else H[i][j] = std::conj(A[j][i]);
            Q[i][j] = (i == j) ? 1.0 : 0.0;
        }
    for (int sweep = 0; sweep < 64; sweep++) {
        double off = abs2(H[0][1]) + abs2(H[0][2]) + abs2(H[1][2]);
        double dia = H[0][0].real() * H[0][0].real() + H[1][1].real() * H[1][1].real() +
                     H[2][2].real() * H[2][2].real();
        if (off <= 1e-36 * dia || off == 0.0) break;
        for (int p = 0; p < 2; p++)
            for (int q = p + 1; q < 3; q++) {
                const cd hpq = H[p][q];
                const double ahpq = std::abs(hpq);
                if (ahpq == 0.0) continue;
                const cd ph = hpq / ahpq;                       // e^{i arg h_pq}
                const double hpp = H[p][p].real(), hqq = H[q][q].real();
                const double tau = (hqq - hpp) / (2.0 * ahpq);
                const double t = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
                const double cth = 1.0 / std::sqrt(1.0 + t * t), sth = t * cth;
                // G = rotation in (p,q) plane with phase: columns p,q of Q
                // new_p = c*col_p - s*conj(ph)*col_q ; new_q = s*ph*col_p + c*col_q
                for (int k = 0; k < 3; k++) {
                    const cd hkp = H[k][p], hkq = H[k][q];
                    H[k][p] = cth * hkp - sth * std::conj(ph) * hkq;
                    H[k][q] = sth * ph * hkp + cth * hkq;
                }
                for (int k = 0; k < 3; k++) {
                    const cd hpk = H[p][k], hqk = H[q][k];
                    H[p][k] = cth * hpk - sth * ph * hqk;
                    H[q][k] = sth * std::conj(ph) * hpk + cth * hqk;
                }
                for (int k = 0; k < 3; k++) {
                    const cd qkp = Q[k][p], qkq = Q[k][q];
                    Q[k][p] = cth * qkp - sth * std::conj(ph) * qkq;
                    Q[k][q] = sth * ph * qkp + cth * qkq;
                }
            }
    }
    for (int i = 0; i < 3; i++) w[i] = H[i][i].real();
}

// Eigenvectors from cross products of the rows of (A - w) (paper, Sec. 2.2),
// with the same fall-back test as function_zheevh3 (su3_matrix.h:298-343):
// fall back when |v|^2 <= 256 eps u^2, u = max|w| (squared when >= 1).
void hybrid_eigen(const cd A[3][3], cd Q[3][3], double w[3]) {
    cardano_eigenvalues(A, w);
    double t = std::max(std::fabs(w[0]), std::max(std::fabs(w[1]), std::fabs(w[2])));
    const double u = (t < 1.0) ? t : t * t;
    const double err = 256.0 * DBL_EPSILON * u * u;
    const double a = A[0][0].real(), b = A[1][1].real();
    // parts of (A - w e1) x (A - w e2) that do not depend on w
    const cd x0 = A[0][1] * A[1][2] - A[0][2] * b;
    const cd x1 = A[0][2] * std::conj(A[0][1]) - A[1][2] * a;
    const double x2 = abs2(A[0][1]);
    for (int col = 0; col < 2; col++) {
        const double wk = w[col];
        const cd v0 = x0 + A[0][2] * wk;
        const cd v1 = x1 + A[1][2] * wk;
        const double v2 = (a - wk) * (b - wk) - x2;
        double nrm = abs2(v0) + abs2(v1) + v2 * v2;
        if (nrm <= err) { jacobi_eigen(A, Q, w); return; }
        nrm = std::sqrt(1.0 / nrm);
        Q[0][col] = v0 * nrm;
        Q[1][col] = v1 * nrm;
        Q[2][col] = v2 * nrm;
    }
    // third eigenvector: conj(v0 x v1)
    Q[0][2] = std::conj(Q[1][0] * Q[2][1] - Q[2][0] * Q[1][1]);
    Q[1][2] = std::conj(Q[2][0] * Q[0][1] - Q[0][0] * Q[2][1]);
    Q[2][2] = std::conj(Q[0][0] * Q[1][1] - Q[1][0] * Q[0][1]);
}

// su3_matrix::exponentiate(sign) (su3_matrix.h:173-253):
//   m <- Q diag(exp(i sign w)) Q^dagger, (Q,w) the eigensystem of m's upper triangle.
M3 exponentiate(const M3& in, double sign) {
    cd A[3][3], Q[3][3];
    double w[3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) A[i][j] = in.m[3 * i + j];
    hybrid_eigen(A, Q, w);
    cd e[3];
    for (int k = 0; k < 3; k++) e[k] = std::exp(cd(0.0, sign * w[k]));
    M3 out;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            cd s = 0.0;
            for (int k = 0; k < 3; k++) s += Q[i][k] * e[k] * std::conj(Q[j][k]);
            out.m[3 * i + j] = s;
        }
    return out;
}

// ---------------------------------------------------------------------------
// Lattice momenta, momenta::set (momenta.h:90-125), single rank (pos = 0).
// ---------------------------------------------------------------------------
void momenta_tables(int N, std::vector<double>& phat2, std::vector<double>& pbarx,
                    std::vector<double>& pbary) {
    const size_t V = (size_t)N * N;
    phat2.resize(V); pbarx.resize(V); pbary.resize(V);
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) {
            double sx = std::sin(M_PI * i / (1.0 * N));
            double sy = std::sin(M_PI * j / (1.0 * N));
            phat2[(size_t)i * N + j] = 4.0 * std::pow(sx, 2.0) + 4.0 * std::pow(sy, 2.0);
            pbarx[(size_t)i * N + j] = std::sin(2.0 * M_PI * i / (1.0 * N));
            pbary[(size_t)i * N + j] = std::sin(2.0 * M_PI * j / (1.0 * N));
        }
}

}  // namespace

// Running-coupling parameters.  The reference hard-codes them (field.h:2046,
// 2257-2273): alpha(q2) = (4 pi / beta0) / ln[ ((mu0/L)^2)^(1/c) + (q2 N^2 / (L N)^2)^(1/c) ]^c
// with beta0 = 11 - 2*3/3 = 9, mu0/Lambda = 15/6, Lambda*L = 6, c = 0.2.
extern "C" struct or_alpha {
    double lambda_L;         // Lambda * N (reference: 6)
    double mu0_over_lambda;  // reference: 2.5
    double freeze_c;         // reference: 0.2
    double beta0;            // reference: 9
};

namespace {

or_alpha alpha_default() { return or_alpha{6.0, 2.5, 0.2, 9.0}; }

// sqrt(alpha) in the operation order of field.h:2257-2273 (noise variant)
double sqrt_alpha_noise(double q2, int N, const or_alpha& a) {
    const double tmp = std::pow(a.mu0_over_lambda * a.mu0_over_lambda, 1.0 / a.freeze_c);
    const double tmp2 = 4.0 * M_PI / a.beta0;
    return std::sqrt(tmp2 / std::log(std::pow(tmp + std::pow((q2 * N * N) / a.lambda_L / a.lambda_L,
                                                              1.0 / a.freeze_c), a.freeze_c)));
}

// alpha in the operation order of field.h:2045-2046 / 2032-2033 (kernel variant)
double alpha_kernel(double q2, int N, const or_alpha& a) {
    return 4.0 * M_PI / (a.beta0 * std::log(std::pow(
        std::pow(a.mu0_over_lambda * a.mu0_over_lambda, 1.0 / a.freeze_c) +
        std::pow((q2 * N * N) / a.lambda_L / a.lambda_L, 1.0 / a.freeze_c), a.freeze_c)));
}

// ---------------------------------------------------------------------------
// Unitary 2-D DFT of each of t interleaved components, fftw2D::execute2D
// (mpi_fftw_class.h:191-315): dir = 1 -> exp(-i k.x) (FFTW_FORWARD), dir = 0 ->
// exp(+i k.x) (FFTW_BACKWARD); both scaled by 1/sqrt(N*N) afterwards.
// Iterative radix-2 transform with twiddles rounded from long double.
// ---------------------------------------------------------------------------
struct Plan1D {
    int n = 0, logn = 0;
    std::vector<cd> tw;        // exp(-2 pi i m / n), m < n/2
    std::vector<int> rev;
};

void plan_init(Plan1D& p, int n) {
    p.n = n;
    p.logn = 0;
    while ((1 << p.logn) < n) p.logn++;
    p.tw.resize(n / 2);
    for (int m = 0; m < n / 2; m++) {
        long double ang = -2.0L * 3.141592653589793238462643383279502884L * (long double)m / (long double)n;
        p.tw[m] = cd((double)cosl(ang), (double)sinl(ang));
    }
    p.rev.resize(n);
    for (int i = 0; i < n; i++) {
        int r = 0;
        for (int b = 0; b < p.logn; b++)
            if (i & (1 << b)) r |= 1 << (p.logn - 1 - b);
        p.rev[i] = r;
    }
}

void fft1d(const Plan1D& p, cd* x, bool forward) {
    const int n = p.n;
    for (int i = 0; i < n; i++)
        if (i < p.rev[i]) std::swap(x[i], x[p.rev[i]]);
    for (int len = 2; len <= n; len <<= 1) {
        const int half = len >> 1, step = n / len;
        for (int s = 0; s < n; s += len)
            for (int k = 0; k < half; k++) {
                cd w = p.tw[k * step];
                if (!forward) w = std::conj(w);
                const cd a = x[s + k], b = x[s + k + half] * w;
                x[s + k] = a + b;
                x[s + k + half] = a - b;
            }
    }
}

void fft2d_plane(int N, cd* plane, bool forward, const Plan1D& p) {
    // rows (y contiguous)
#pragma omp parallel for schedule(static)
    for (int x = 0; x < N; x++) fft1d(p, plane + (size_t)x * N, forward);
    // columns
#pragma omp parallel
    {
        std::vector<cd> col(N);
#pragma omp for schedule(static)
        for (int y = 0; y < N; y++) {
            for (int x = 0; x < N; x++) col[x] = plane[(size_t)x * N + y];
            fft1d(p, col.data(), forward);
            for (int x = 0; x < N; x++) plane[(size_t)x * N + y] = col[x];
        }
    }
}

// lambda-basis map used by setMVModel (field.h:714-745) and generate_gaussian
// (field.h:2169-2231): the code adds, per site,
//   u1 += n0, u3 += n0;  u1 += i n1, u3 -= i n1;  u0 += n2, u4 -= n2;
//   u2 += n3, u6 += n3;  u2 += i n4, u6 -= i n4;  u5 += n5, u7 += n5;
//   u5 += i n6, u7 -= i n6;  u0,u4 += n7/sqrt3, u8 += -2 n7/sqrt3
// i.e. the generators {l1,-l2,l3,l4,-l5,l6,-l7,l8}.
void lambda_site(const double n[8], cd* u) {
    for (int j = 0; j < 9; j++) u[j] = 0.0;
    const cd I(0.0, 1.0);
    const double s3 = std::sqrt(3.0);
    u[1] += n[0]; u[3] += n[0];
    u[1] += I * n[1]; u[3] -= I * n[1];
    u[0] += n[2]; u[4] -= n[2];
    u[2] += n[3]; u[6] += n[3];
    u[2] += I * n[4]; u[6] -= I * n[4];
    u[5] += n[5]; u[7] += n[5];
    u[5] += I * n[6]; u[7] -= I * n[6];
    u[0] += n[7] / s3; u[4] += n[7] / s3; u[8] += -2.0 * n[7] / s3;
}

struct FFTCache {
    int N = 0;
    Plan1D p;
};
thread_local FFTCache g_fft;

const Plan1D& plan_for(int N) {
    if (g_fft.N != N) { plan_init(g_fft.p, N); g_fft.N = N; }
    return g_fft.p;
}

}  // namespace

extern "C" {

// ---- version / sanity -------------------------------------------------------
int or_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void or_set_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

// 1: uxiulocal and update_uf run on one thread, as the reference ships them
// (their OpenMP pragmas are commented out, field.h:1901, 1974); 0: all site
// loops use every thread.  For the CPU baseline's two variants.
static int g_as_shipped = 0;
void or_set_as_shipped(int on) { g_as_shipped = on ? 1 : 0; }

// ---- SU(3) ------------------------------------------------------------------
// su3_matrix::exponentiate(sign), su3_matrix.h:173-253.  in/out: 18 doubles.
int or_su3_exponentiate(const double* in18, double sign, double* out18) {
    M3 a;
    for (int k = 0; k < 9; k++) a.m[k] = C(in18)[k];
    M3 r = exponentiate(a, sign);
    for (int k = 0; k < 9; k++) C(out18)[k] = r.m[k];
    return 0;
}

// function_zheevh3 (su3_matrix.h:256-359): eigenvalues w[3], eigenvectors in
// the columns of Q (row-major 18 doubles).
int or_su3_eigen(const double* in18, double* Q18, double* w3) {
    cd A[3][3], Q[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) A[i][j] = C(in18)[3 * i + j];
    hybrid_eigen(A, Q, w3);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) C(Q18)[3 * i + j] = Q[i][j];
    return 0;
}

// su3_matrix::operator* (su3_matrix.h:155-168)
int or_su3_mul(const double* a18, const double* b18, double* c18) {
    M3 a, b;
    for (int k = 0; k < 9; k++) { a.m[k] = C(a18)[k]; b.m[k] = C(b18)[k]; }
    M3 c = mul(a, b);
    for (int k = 0; k < 9; k++) C(c18)[k] = c.m[k];
    return 0;
}

// ---- tables -----------------------------------------------------------------
// momenta::set (momenta.h:90-125)
int or_momenta(int N, double* phat2, double* pbarx, double* pbary) {
    if (!pow2(N)) return -1;
    std::vector<double> a, b, c;
    momenta_tables(N, a, b, c);
    std::memcpy(phat2, a.data(), a.size() * sizeof(double));
    std::memcpy(pbarx, b.data(), b.size() * sizeof(double));
    std::memcpy(pbary, c.data(), c.size() * sizeof(double));
    return 0;
}

// sqrt(alpha(q2)) of generate_gaussian_with_noise_coupling_constant (field.h:2257-2273)
double or_sqrt_alpha_noise(double q2, int N, const or_alpha* a) {
    return sqrt_alpha_noise(q2, N, a ? *a : alpha_default());
}

// alpha(q2) of prepare_A_local's SQRT branch (field.h:2045-2046)
double or_alpha_kernel(double q2, int N, const or_alpha* a) {
    return alpha_kernel(q2, N, a ? *a : alpha_default());
}

// ---- FFT ----------------------------------------------------------------------
// fftw2D::execute2D(lfield<T,t>*, dir), mpi_fftw_class.h:191-315
int or_fft2d(int N, int t, double* f, int dir) {
    if (!pow2(N) || t < 1) return -1;
    const size_t V = (size_t)N * N;
    const Plan1D& p = plan_for(N);
    std::vector<cd> plane(V);
    const double scale = 1.0 / std::sqrt((double)((long)N * N));
    cd* u = C(f);
    for (int k = 0; k < t; k++) {
        for (size_t i = 0; i < V; i++) plane[i] = u[i * t + k];
        fft2d_plane(N, plane.data(), dir != 0, p);
        for (size_t i = 0; i < V; i++)
            u[i * t + k] = cd(plane[i].real() * scale, plane[i].imag() * scale);
    }
    return 0;
}

// Plain O(V^2) DFT with the same conventions, for checking or_fft2d on tiny N.
int or_dft2d_bruteforce(int N, int t, double* f, int dir) {
    if (N < 1 || t < 1) return -1;
    const size_t V = (size_t)N * N;
    std::vector<cd> out(V * t);
    const long double twopi = 6.283185307179586476925286766559005768L;
    const double sgn = dir ? -1.0 : 1.0;
    cd* u = C(f);
    for (int kx = 0; kx < N; kx++)
        for (int ky = 0; ky < N; ky++)
            for (int k = 0; k < t; k++) {
                long double sr = 0, si = 0;
                for (int x = 0; x < N; x++)
                    for (int y = 0; y < N; y++) {
                        long double ang = sgn * twopi * (long double)((kx * x + ky * y) % N) / N;
                        long double c = cosl(ang), s = sinl(ang);
                        cd v = u[((size_t)x * N + y) * t + k];
                        sr += v.real() * c - v.imag() * s;
                        si += v.real() * s + v.imag() * c;
                    }
                out[((size_t)kx * N + ky) * t + k] = cd((double)(sr / N), (double)(si / N));
            }
    std::memcpy(f, out.data(), V * t * sizeof(cd));
    return 0;
}

// ---- fields -------------------------------------------------------------------
// lfield::setToUnit (field.h:196-211)
int or_set_unit(int N, double* U) {
    const size_t V = (size_t)N * N;
    cd* u = C(U);
    for (size_t i = 0; i < V; i++)
        for (int k = 0; k < 9; k++) u[i * 9 + k] = (k == 0 || k == 4 || k == 8) ? 1.0 : 0.0;
    return 0;
}

// generate_gaussian (field.h:2116-2243) with injected normals.
// z: 16 per site in draw order n0,m0,n1,m1,... (field.h:2162-2165).
int or_generate_gaussian(int N, const double* z, double* xix, double* xiy) {
    const size_t V = (size_t)N * N;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < V; i++) {
        double n[8], m[8];
        for (int k = 0; k < 8; k++) { n[k] = z[16 * i + 2 * k]; m[k] = z[16 * i + 2 * k + 1]; }
        lambda_site(n, C(xix) + 9 * i);
        lambda_site(m, C(xiy) + 9 * i);
    }
    return 0;
}

// generate_gaussian_with_noise_coupling_constant (field.h:2245-2365): the normals
// are multiplied by sqrt(alpha(phat2(i))) with i the POSITION index (field.h:2273).
int or_generate_gaussian_noise_coupling(int N, const double* z, const or_alpha* ap,
                                        double* xix, double* xiy) {
    if (!pow2(N)) return -1;
    const or_alpha a = ap ? *ap : alpha_default();
    const size_t V = (size_t)N * N;
    std::vector<double> ph, bx, by;
    momenta_tables(N, ph, bx, by);
    for (size_t i = 0; i < V; i++) {
        const double s = sqrt_alpha_noise(ph[i], N, a);
        double n[8], m[8];
        for (int k = 0; k < 8; k++) { n[k] = z[16 * i + 2 * k] * s; m[k] = z[16 * i + 2 * k + 1] * s; }
        lambda_site(n, C(xix) + 9 * i);
        lambda_site(m, C(xiy) + 9 * i);
    }
    return 0;
}

// setMVModel (field.h:683-757) with injected standard normals z (8 per site);
// the reference draws N(0, sigma) with sigma = g mu / sqrt(Ny) (field.h:697),
// which libstdc++ forms as z*sigma + 0.
int or_set_mv_model(int N, const double* z, double sigma, double* f) {
    const size_t V = (size_t)N * N;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < V; i++) {
        double n[8];
        for (int k = 0; k < 8; k++) n[k] = z[8 * i + k] * sigma + 0.0;
        lambda_site(n, C(f) + 9 * i);
    }
    return 0;
}

// solvePoisson (field.h:867-877): f(k) *= -g / (-phat2(k) + mass^2)
int or_solve_poisson(int N, double* f, double mass, double g) {
    if (!pow2(N)) return -1;
    std::vector<double> ph, bx, by;
    momenta_tables(N, ph, bx, by);
    const size_t V = (size_t)N * N;
    cd* u = C(f);
    for (size_t i = 0; i < V; i++) {
        const cd fac(-1.0 * g / (-ph[i] + mass * mass), 0.0);
        for (int k = 0; k < 9; k++) u[i * 9 + k] *= fac;
    }
    return 0;
}

// lfield::operator*= (field.h:337-365): U <- U * exp(+i f), f read through its
// upper triangle.
int or_mul_exp(int N, double* U, const double* f) {
    const size_t V = (size_t)N * N;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < V; i++) {
        M3 A, B;
        for (int k = 0; k < 9; k++) { A.m[k] = C(U)[9 * i + k]; B.m[k] = C(f)[9 * i + k]; }
        B = exponentiate(B, 1.0);
        M3 c = mul(A, B);
        for (int k = 0; k < 9; k++) C(U)[9 * i + k] = c.m[k];
    }
    return 0;
}

// prepare_A_local, 7-argument form (field.h:1998-2075).  Fields are in k-space.
// coupling: 0 SQRT (c = alpha(k)), 1 NOISE (c = 1), 3 FIXED (c = 1),
//           4 CORRELATED (c = 1; the noise carries sqrt(alpha));
//           2 HATTA leaves c uninitialised in the reference and is rejected.
// kernel: 0 LINEAR, 1 SIN.
int or_prepare_A_local(int N, double* A, const double* xix, const double* xiy,
                       int coupling, int kernel, const or_alpha* ap) {
    if (!pow2(N) || coupling == 2 || coupling < 0 || coupling > 4 || kernel < 0 || kernel > 1)
        return -1;
    const or_alpha a = ap ? *ap : alpha_default();
    std::vector<double> ph, bx, by;
    momenta_tables(N, ph, bx, by);
#pragma omp parallel for collapse(2) schedule(static)
    for (int ix = 0; ix < N; ix++)
        for (int iy = 0; iy < N; iy++) {
            const size_t i = (size_t)ix * N + iy;
            int xx = ix, yy = iy;
            if (xx >= N / 2) xx -= N;
            if (yy >= N / 2) yy -= N;
            const double px = 2.0 * M_PI * xx / (1.0 * N);
            const double py = 2.0 * M_PI * yy / (1.0 * N);
            cd E[9];
            for (int k = 0; k < 9; k++) E[k] = 0.0;
            if (std::fabs(px * px + py * py) > 10e-9) {
                double c = 1.0;
                cd KX, KY;
                if (kernel == 0) {
                    if (coupling == 0) c = alpha_kernel(px * px + py * py, N, a);
                    KX = cd(0.0, -2.0 * M_PI * std::sqrt(c) * px / (px * px + py * py));
                    KY = cd(0.0, -2.0 * M_PI * std::sqrt(c) * py / (px * px + py * py));
                } else {
                    if (coupling == 0) c = alpha_kernel(ph[i], N, a);
                    KX = cd(0.0, -2.0 * M_PI * std::sqrt(c) * bx[i] / ph[i]);
                    KY = cd(0.0, -2.0 * M_PI * std::sqrt(c) * by[i] / ph[i]);
                }
                for (int k = 0; k < 9; k++)
                    E[k] = KX * C(xix)[9 * i + k] + KY * C(xiy)[9 * i + k];
            }
            for (int k = 0; k < 9; k++) C(A)[9 * i + k] = E[k];
        }
    return 0;
}

// uxiulocal (field.h:1892-1933): V_i = (U xi_i) U^dagger
int or_uxiulocal(int N, double* Vx, double* Vy, const double* U, const double* xix,
                 const double* xiy) {
    const size_t V = (size_t)N * N;
#pragma omp parallel for schedule(static) if (!g_as_shipped)
    for (size_t i = 0; i < V; i++) {
        M3 u, bx, by;
        for (int k = 0; k < 9; k++) {
            u.m[k] = C(U)[9 * i + k];
            bx.m[k] = C(xix)[9 * i + k];
            by.m[k] = C(xiy)[9 * i + k];
        }
        M3 d = dagger(u);
        M3 ex = mul(mul(u, bx), d), ey = mul(mul(u, by), d);
        for (int k = 0; k < 9; k++) { C(Vx)[9 * i + k] = ex.m[k]; C(Vy)[9 * i + k] = ey.m[k]; }
    }
    return 0;
}

// update_uf (field.h:1966-1996): U <- exp(-i(-sqrt(s) B)) U exp(-i(sqrt(s) A))
int or_update_uf(int N, double* U, const double* B, const double* A, double step) {
    const size_t V = (size_t)N * N;
    const double ss = std::sqrt(step);
#pragma omp parallel for schedule(static) if (!g_as_shipped)
    for (size_t i = 0; i < V; i++) {
        M3 u, b, a;
        for (int k = 0; k < 9; k++) {
            u.m[k] = C(U)[9 * i + k];
            b.m[k] = -ss * C(B)[9 * i + k];
            a.m[k] = ss * C(A)[9 * i + k];
        }
        b = exponentiate(b, -1.0);
        a = exponentiate(a, -1.0);
        M3 e = mul(mul(b, u), a);
        for (int k = 0; k < 9; k++) C(U)[9 * i + k] = e.m[k];
    }
    return 0;
}

// lfield::trace (field.h:1375-1432): corr(k) = tr[(U/V)(U/V)^dagger]; corr is an
// lfield<double,1> (complex, 2 doubles per site).
int or_trace(int N, const double* U, double* corr) {
    const size_t V = (size_t)N * N;
    const double vol = 1.0 * N * N;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < V; i++) {
        M3 a;
        for (int k = 0; k < 9; k++) a.m[k] = C(U)[9 * i + k] / vol;
        M3 b = dagger(a);
        M3 c = mul(a, b);
        C(corr)[i] = c.m[0] + c.m[4] + c.m[8];
    }
    return 0;
}

// gfield::average_and_symmetrize (field.h:1460-1492) on an N x N complex t=1 field
int or_average_and_symmetrize(int N, double* corr) {
    const size_t V = (size_t)N * N;
    std::vector<cd> tmp(V);
    cd* u = C(corr);
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++)
            tmp[(size_t)i * N + j] = 0.5 * (u[(size_t)i * N + j] + u[(size_t)j * N + i]);
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) {
            const int a = std::abs(i - N) % N, b = std::abs(j - N) % N;
            u[(size_t)i * N + j] = 0.25 * (tmp[(size_t)i * N + j] + tmp[(size_t)a * N + j] +
                                           tmp[(size_t)i * N + b] + tmp[(size_t)a * N + b]);
        }
    return 0;
}

// gfield::reduce(sum, err, mpi) (field.h:1521-1536): sum += c, err += c^2
// (std::pow(complex, 2.0) of a real positive value is its real square).
int or_reduce(int N, const double* corr, double* sum, double* err) {
    const size_t V = (size_t)N * N;
    for (size_t i = 0; i < V; i++) {
        const cd c = C(corr)[i];
        C(sum)[i] += c;
        C(err)[i] += std::pow(c, 2.0);
    }
    return 0;
}

// ---- composed driver pieces -------------------------------------------------
// Per-mode correlated noise (coupling 4), in momentum space: each of the t
// components of the spectrum f is multiplied by sqrt(alpha(phat2(k))), alpha
// in prepare_A_local's operation order (field.h:2045-2046).  With xi~ =
// F^-1[sqrt(alpha) F xi] the noise covariance is the circulant
// C(x - y) = (1/V) sum_k alpha(k) e^{ik(x-y)}: the matrix whose dense Cholesky
// factor the reference's POSITION_EVOLUTION path applies (field.h:1633-1651
// builds alpha(k)/sqrt(V), main_optimized.cpp:145-157 transforms it backward
// and factors it with matrix.h:151-200).  The per-mode factor sqrt(alpha) is the
// Fourier-diagonal square root of that matrix.
int or_scale_by_sqrt_alpha(int N, int t, double* f, const or_alpha* ap) {
    if (!pow2(N) || t < 1) return -1;
    const or_alpha a = ap ? *ap : alpha_default();
    const size_t V = (size_t)N * N;
    std::vector<double> ph, bx, by;
    momenta_tables(N, ph, bx, by);
    for (size_t i = 0; i < V; i++) {
        const double s = std::sqrt(alpha_kernel(ph[i], N, a));
        for (int k = 0; k < 2 * t; k++) f[2 * t * i + k] *= s;
    }
    return 0;
}

// xi~ = F^-1[sqrt(alpha(k)) F xi] on t-component position-space fields, in place
// (the map of or_langevin_step's coupling 4, exposed for the covariance pin).
int or_correlate_noise(int N, int t, double* f, const or_alpha* ap) {
    if (or_fft2d(N, t, f, 1)) return -1;
    if (or_scale_by_sqrt_alpha(N, t, f, ap)) return -1;
    return or_fft2d(N, t, f, 0);
}

// One Langevin step of main_optimized's MOMENTUM branch (main_optimized.cpp:243-304)
// with injected normals z (16 per site).  coupling as in or_prepare_A_local;
// NOISE uses the position-index noise scaling (field.h:2273);
// CORRELATED multiplies F xi by sqrt(alpha(phat2(k))) (or_scale_by_sqrt_alpha)
// before both the kernel (factor 1) and the inverse transform.
int or_langevin_step(int N, double* U, const double* z, int coupling, int kernel, double step,
                     const or_alpha* ap) {
    if (!pow2(N) || coupling == 2 || coupling < 0 || coupling > 4) return -1;
    const or_alpha a = ap ? *ap : alpha_default();
    const size_t V = (size_t)N * N;
    std::vector<double> xix(18 * V), xiy(18 * V), A(18 * V), B(18 * V), Vx(18 * V), Vy(18 * V);
    if (coupling == 1) or_generate_gaussian_noise_coupling(N, z, &a, xix.data(), xiy.data());
    else or_generate_gaussian(N, z, xix.data(), xiy.data());
    or_fft2d(N, 9, xix.data(), 1);
    or_fft2d(N, 9, xiy.data(), 1);
    if (coupling == 4) {
        or_scale_by_sqrt_alpha(N, 9, xix.data(), &a);
        or_scale_by_sqrt_alpha(N, 9, xiy.data(), &a);
    }
    or_prepare_A_local(N, A.data(), xix.data(), xiy.data(), coupling, kernel, &a);
    or_fft2d(N, 9, A.data(), 0);
    or_fft2d(N, 9, xix.data(), 0);
    or_fft2d(N, 9, xiy.data(), 0);
    or_uxiulocal(N, Vx.data(), Vy.data(), U, xix.data(), xiy.data());
    or_fft2d(N, 9, Vx.data(), 1);
    or_fft2d(N, 9, Vy.data(), 1);
    or_prepare_A_local(N, B.data(), Vx.data(), Vy.data(), coupling, kernel, &a);
    or_fft2d(N, 9, B.data(), 0);
    or_update_uf(N, U, B.data(), A.data(), step);
    return 0;
}

// MV initial condition, main_optimized.cpp:178-191: U = 1; per sheet
// f = setMVModel; F f; solvePoisson(mass*g^2*mu, g); F^-1 f; U *= f.
// z: Ny * V * 8 standard normals (sheet-major).  mu is MV_class's mu_parameter.
// zero_mode = 1: each sheet's k = 0 mode is set to zero after solvePoisson
// instead of carrying the factor -g/m^2 (the library's projection option; the
// reference has no such switch)
int or_mv_init_opt(int N, double* U, const double* z, int Ny, double g, double mu, double mass, int zero_mode) {
    if (!pow2(N) || Ny < 1) return -1;
    const size_t V = (size_t)N * N;
    std::vector<double> f(18 * V);
    or_set_unit(N, U);
    const double sigma = g * mu / std::sqrt((double)Ny);
    for (int s = 0; s < Ny; s++) {
        or_set_mv_model(N, z + (size_t)s * V * 8, sigma, f.data());
        or_fft2d(N, 9, f.data(), 1);
        or_solve_poisson(N, f.data(), mass * std::pow(g, 2.0) * mu, g);
        if (zero_mode)
            for (int k = 0; k < 18; k++) f[k] = 0.0;
        or_fft2d(N, 9, f.data(), 0);
        or_mul_exp(N, U, f.data());
    }
    return 0;
}

int or_mv_init(int N, double* U, const double* z, int Ny, double g, double mu, double mass) {
    return or_mv_init_opt(N, U, z, Ny, g, mu, mass, 0);
}

// Measurement, main_optimized.cpp:321-333 (single rank): copy, F U, trace,
// allgather (identity), average_and_symmetrize.  corr: complex t=1 field.
int or_measure_corr(int N, const double* U, double* corr) {
    const size_t V = (size_t)N * N;
    std::vector<double> uc(U, U + 18 * V);
    or_fft2d(N, 9, uc.data(), 1);
    or_trace(N, uc.data(), corr);
    return or_average_and_symmetrize(N, corr);
}

// ---- POSITION_EVOLUTION (main_optimized.cpp:271-297, field.h:2368-2519) ----
// Kernel weights of the position-space step for every displacement
// r = (rx, ry) = (x - xx, y - yy) mod N, at index rx N + ry:
//   SIN kernel (positions.h:131-141 tables): dx = 0.5 N sin(2 pi rx/N)/pi,
//     dy likewise, rrr = pow(N sin(pi rx/N)/pi, 2) + pow(N sin(pi ry/N)/pi, 2);
//   LINEAR kernel: dx, dy = rx, ry wrapped into [-N/2, N/2), rrr = dx^2 + dy^2;
//   SQRT coupling (0): c = sqrt(4 pi / ((11 - 2*3/3) log(pow(lambda + 1.26 /
//     pow(36 rrr/N^2, 1/0.2), 0.2)))) with lambda = pow(225/36, 1/0.2)
//     (field.h:2449-2456); NOISE (1) and FIXED (3): c = 1 (the reference applies
//     its Cholesky factor to the noise of NOISE beforehand, main_optimized.cpp:277-282);
//   wa = c dx / rrr, wb = c dy / rrr where rrr > 10e-9, else 0 (field.h:2460-2468).
int or_position_weights(int N, int coupling, int kernel, double* wa, double* wb) {
    if (!pow2(N) || (coupling != 0 && coupling != 1 && coupling != 3) || (kernel != 0 && kernel != 1)) return -1;
    const double lambda = std::pow(15.0 * 15.0 / 6.0 / 6.0, 1.0 / 0.2);
    for (int rx = 0; rx < N; rx++)
        for (int ry = 0; ry < N; ry++) {
            double dx, dy, rrr;
            if (kernel == 1) {
                dx = 0.5 * N * std::sin(2.0 * M_PI * rx / (1.0 * N)) / M_PI;
                dy = 0.5 * N * std::sin(2.0 * M_PI * ry / (1.0 * N)) / M_PI;
                const double bx = N * std::sin(M_PI * rx / (1.0 * N)) / M_PI;
                const double by = N * std::sin(M_PI * ry / (1.0 * N)) / M_PI;
                rrr = std::pow(bx, 2.0) + std::pow(by, 2.0);
            } else {
                dx = rx >= N / 2 ? rx - N : rx;
                dy = ry >= N / 2 ? ry - N : ry;
                rrr = 1.0 * (dx * dx + dy * dy);
            }
            double c = 1.0;
            if (coupling == 0)
                c = std::sqrt(4.0 * M_PI /
                              ((11.0 - 2.0 * 3.0 / 3.0) *
                               std::log(std::pow(lambda + 1.26 / std::pow(6.0 * 6.0 * rrr / N / N, 1.0 / 0.2), 0.2))));
            const size_t i = (size_t)rx * N + ry;
            wa[i] = rrr > 10e-9 ? c * dx / rrr : 0.0;
            wb[i] = rrr > 10e-9 ? c * dy / rrr : 0.0;
        }
    return 0;
}

// prepare_A_and_B_local for every site (the loop of main_optimized.cpp:288-297):
//   A(x) = sum_y E_xy,  B(x) = sum_y U(y) E_xy U(y)^dagger,
//   E_xy = wa(x - y) xi_x(y) + wb(x - y) xi_y(y)
// with the products in the reference's order (field.h:2478-2506): C = A xi_x,
// D = B xi_y entry by entry, E = C + D, K = (G E) H with H = G^dagger; sums
// over y in row-major order (the reference's OpenMP reduction has no fixed order).
int or_prepare_A_and_B_local(int N, double* A, double* B, const double* xix, const double* xiy, const double* U,
                             int coupling, int kernel) {
    const size_t V = (size_t)N * N;
    std::vector<double> wa(V), wb(V);
    if (or_position_weights(N, coupling, kernel, wa.data(), wb.data())) return -1;
    const cd* X = C(xix);
    const cd* Y = C(xiy);
    const cd* G = C(U);
#pragma omp parallel for schedule(dynamic, 16)
    for (size_t t = 0; t < V; t++) {
        const int x = (int)(t / N), y = (int)(t % N);
        cd sa[9], sb[9];
        for (int k = 0; k < 9; k++) sa[k] = sb[k] = 0.0;
        for (int xx = 0; xx < N; xx++)
            for (int yy = 0; yy < N; yy++) {
                const size_t i = (size_t)xx * N + yy;
                const size_t r = (size_t)((x - xx + N) % N) * N + (size_t)((y - yy + N) % N);
                const cd a(wa[r], 0.0), b(wb[r], 0.0);
                M3 E, Gm;
                for (int k = 0; k < 9; k++) {
                    E.m[k] = a * X[9 * i + k] + b * Y[9 * i + k];
                    Gm.m[k] = G[9 * i + k];
                }
                const M3 K = mul(mul(Gm, E), dagger(Gm));
                for (int k = 0; k < 9; k++) {
                    sa[k] += E.m[k];
                    sb[k] += K.m[k];
                }
            }
        for (int k = 0; k < 9; k++) {
            C(A)[9 * t + k] = sa[k];
            C(B)[9 * t + k] = sb[k];
        }
    }
    return 0;
}

// ---- dipole correlator (beyond the reference: derived by Wiener-Khinchin) ----
// S(r) = (V/3) Re sum_k corr(k) exp(+i k.r), corr the (real) output of
// or_measure_corr.  Here by direct O(V^2) summation; for checks at small N.
int or_dipole_from_corr_bruteforce(int N, const double* corr, double* S) {
    const size_t V = (size_t)N * N;
    const long double twopi = 6.283185307179586476925286766559005768L;
    for (int rx = 0; rx < N; rx++)
        for (int ry = 0; ry < N; ry++) {
            long double s = 0;
            for (int kx = 0; kx < N; kx++)
                for (int ky = 0; ky < N; ky++) {
                    long double ang = twopi * (long double)((kx * rx + ky * ry) % N) / N;
                    s += corr[2 * ((size_t)kx * N + ky)] * cosl(ang);
                }
            S[(size_t)rx * N + ry] = (double)(s * (long double)V / 3.0L);
        }
    return 0;
}

// S(r) = (1/3V) sum_x Re tr U(x) U^dagger(x + r), directly in position space.
int or_dipole_position_space(int N, const double* U, double* S) {
    const size_t V = (size_t)N * N;
#pragma omp parallel for collapse(2) schedule(static)
    for (int rx = 0; rx < N; rx++)
        for (int ry = 0; ry < N; ry++) {
            double acc = 0.0;
            for (int x = 0; x < N; x++)
                for (int y = 0; y < N; y++) {
                    const cd* a = C(U) + 9 * ((size_t)x * N + y);
                    const cd* b = C(U) + 9 * ((size_t)((x + rx) % N) * N + (y + ry) % N);
                    for (int k = 0; k < 9; k++) acc += (a[k] * std::conj(b[k])).real();
                }
            S[(size_t)rx * N + ry] = acc / (3.0 * V);
        }
    return 0;
}

// Radial binning of S(r) by minimal-image distance: bin b holds b <= |r| < b+1.
// rbar[b] = mean |r| of its members, Sb[b] = mean S, cnt[b] = members.
int or_radial_bins(int N, const double* S, int nbins, double* rbar, double* Sb, double* cnt) {
    std::vector<long double> sr(nbins, 0), ss(nbins, 0), sc(nbins, 0);
    for (int x = 0; x < N; x++)
        for (int y = 0; y < N; y++) {
            const int dx = x < N / 2 ? x : x - N, dy = y < N / 2 ? y : y - N;
            const double r = std::sqrt((double)(dx * dx + dy * dy));
            const int b = (int)std::floor(r);
            if (b >= nbins) continue;
            sr[b] += r; ss[b] += S[(size_t)x * N + y]; sc[b] += 1;
        }
    for (int b = 0; b < nbins; b++) {
        cnt[b] = (double)sc[b];
        rbar[b] = sc[b] > 0 ? (double)(sr[b] / sc[b]) : 0.0;
        Sb[b] = sc[b] > 0 ? (double)(ss[b] / sc[b]) : 0.0;
    }
    return 0;
}

// ---- Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) ----------------------
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Standard normals of the library's noise stream.  For site s and call j the
// counter is {s, step, tag, j} and the key {seed, traj}; each call gives one
// Box-Muller pair from two 53-bit uniforms in (0,1):
//   u = (((hi << 32) | lo) >> 11 + 0.5) * 2^-53,
//   z0 = sqrt(-2 ln u0) cos(2 pi u1),  z1 = sqrt(-2 ln u0) sin(2 pi u1).
// Output: per site `per_site` normals (even), call j filling z[2j], z[2j+1].
int or_philox_normals(int N, uint32_t seed, uint32_t traj, uint32_t step, uint32_t tag,
                      int per_site, double* z) {
    if (per_site <= 0 || (per_site & 1)) return -1;
    const size_t V = (size_t)N * N;
    const uint32_t key[2] = {seed, traj};
    for (size_t s = 0; s < V; s++)
        for (int j = 0; j < per_site / 2; j++) {
            const uint32_t ctr[4] = {(uint32_t)s, step, tag, (uint32_t)j};
            uint32_t o[4];
            or_philox4x32_10(ctr, key, o);
            const uint64_t a = ((uint64_t)o[0] << 32) | o[1];
            const uint64_t b = ((uint64_t)o[2] << 32) | o[3];
            const double u0 = ((double)(a >> 11) + 0.5) * 0x1.0p-53;
            const double u1 = ((double)(b >> 11) + 0.5) * 0x1.0p-53;
            const double r = std::sqrt(-2.0 * std::log(u0));
            const long double th = 6.283185307179586476925286766559005768L * (long double)u1;
            z[s * per_site + 2 * j] = r * (double)cosl(th);
            z[s * per_site + 2 * j + 1] = r * (double)sinl(th);
        }
    return 0;
}

// Momentum-space Langevin noise of the library (jw_set_noise_space): for mode
// k = kx N + ky and plane pair p the Philox call {k, step, tag, p} (key
// {seed, traj}) gives four 32-bit uniforms u_j = (w_j + 1/2) 2^-32 and two
// Box-Muller pairs
//   (a, b) = sqrt(-2 ln u0) (cos 2 pi u1, sin 2 pi u1),  (c, d) from u2, u3,
// the unit-variance spectra P_x = a + i b, P_y = c + i d of the x- and
// y-direction plane pairs.  In position space
//   w_d(x) = N^-1 sum_k exp(+2 pi i k.x / N) P_d(k)     (unitary inverse DFT),
// so that the library's forward transform of w_d is N P_d; site x gets the
// normals n_2p = Re w_x, n_2p+1 = Im w_x, m_2p = Re w_y, m_2p+1 = Im w_y in the
// draw order n0, m0, n1, m1, ... (z[16 x + 4 p + {0, 2, 1, 3}]).
// spec (optional, 16 V doubles): the drawn spectra, per mode
// (P_x re, P_x im, P_y re, P_y im) for p = 0..3.
int or_kspace_normals(int N, uint32_t seed, uint32_t traj, uint32_t step, uint32_t tag, double* z, double* spec) {
    if (!pow2(N)) return -1;
    const size_t V = (size_t)N * N;
    const uint32_t key[2] = {seed, traj};
    std::vector<cd> w(2 * V);
    const Plan1D& plan = plan_for(N);
    for (int p = 0; p < 4; p++) {
        for (size_t k = 0; k < V; k++) {
            const uint32_t ctr[4] = {(uint32_t)k, step, tag, (uint32_t)p};
            uint32_t o[4];
            or_philox4x32_10(ctr, key, o);
            double g[4];
            for (int h = 0; h < 2; h++) {
                const double u0 = ((double)o[2 * h] + 0.5) * 0x1.0p-32;
                const double u1 = ((double)o[2 * h + 1] + 0.5) * 0x1.0p-32;
                const double r = std::sqrt(-2.0 * std::log(u0));
                const long double th = 6.283185307179586476925286766559005768L * (long double)u1;
                g[2 * h] = r * (double)cosl(th);
                g[2 * h + 1] = r * (double)sinl(th);
            }
            w[k] = cd(g[0], g[1]);
            w[V + k] = cd(g[2], g[3]);
            if (spec)
                for (int j = 0; j < 4; j++) spec[16 * k + 4 * p + j] = g[j];
        }
        for (int d = 0; d < 2; d++) fft2d_plane(N, w.data() + d * V, false, plan);
        const double inv = 1.0 / (double)N;
        for (size_t x = 0; x < V; x++) {
            const cd a = w[x] * inv, b = w[V + x] * inv;
            z[16 * x + 4 * p + 0] = a.real();
            z[16 * x + 4 * p + 2] = a.imag();
            z[16 * x + 4 * p + 1] = b.real();
            z[16 * x + 4 * p + 3] = b.imag();
        }
    }
    return 0;
}

}  // extern "C"
