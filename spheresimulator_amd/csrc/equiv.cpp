// Equivalent-source template of the far-field cell aggregation (the role of the
// FMM's upward pass in STKFMM, BIE/SphereSTKSHOperator.cpp:132-141, :365-369).
//
// A cell is a group of nearby spheres inside a ball of radius rho.  Outside a
// larger ball the Stokes flow of all its point forces (velocity and pressure,
// hence also the traction on any surface) is reproduced by K point forces on
// the sphere of radius re*rho around its centre: they are fitted, in the least-
// squares sense, to the velocity the cell's sources induce at Kc check points
// on the sphere of radius rc*rho.  For the unit cell the fit is the matrix
//   P = pinv(G),  G[3i+a][3k+b] = G_ab(rc chk_i - re eq_k)  (the Stokeslet
// (1/8pi)(delta_ab/r + r_a r_b/r^3)), 3K x 3Kc;
// for a cell of radius rho the Stokeslet scales as 1/rho, so the forces are
// rho P u_check.  P comes from a Householder QR of G (cond(G) ~ 1e7 for the
// radii below, well inside FP64).  Both point sets are Fibonacci spirals
// (nearly uniform; the Gauss-Legendre grids cluster at the poles and fit an
// order of magnitude worse).  With K = 200, Kc = 400, re = 1.1, rc = 2.5,
// 16 spheres of radius rho/8 inside the cell and targets at 4 rho, velocity
// and traction are reproduced to about 1e-7 and 1.2e-6 (relative).  Smaller
// templates (Kc = 2K, same radii) trade accuracy for pair evaluations.
#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <thread>

namespace sphstk {

namespace {

void fibonacci_sphere(int K, std::vector<double> &p) {
    p.resize(3 * (size_t)K);
    const double golden = kPi * (1.0 + std::sqrt(5.0));
    for (int i = 0; i < K; i++) {
        const double z = 1.0 - 2.0 * (i + 0.5) / K, s = std::sqrt(std::max(0.0, 1.0 - z * z)), ph = golden * (i + 0.5);
        p[3 * i] = s * std::cos(ph);
        p[3 * i + 1] = s * std::sin(ph);
        p[3 * i + 2] = z;
    }
}

// Runs body(j) for j in [0, n) on up to 16 host threads.
template <typename F>
void for_columns(int n, F body) {
    const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (nt == 1 || n < 64) {
        for (int j = 0; j < n; j++) body(j);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < nt; t++)
        th.emplace_back([&, t] {
            for (int j = t; j < n; j += nt) body(j);
        });
    for (auto &x : th) x.join();
}

EquivTables build(int K) {
    EquivTables T;
    T.K = K;
    T.Kc = 2 * K;
    T.re = 1.1;
    T.rc = 2.5;
    fibonacci_sphere(T.K, T.eq);
    fibonacci_sphere(T.Kc, T.chk);
    const int m = 3 * T.Kc, n = 3 * T.K;
    // G, column-major (m x n)
    std::vector<double> A((size_t)m * n);
    for (int i = 0; i < T.Kc; i++)
        for (int k = 0; k < T.K; k++) {
            double r[3], r2 = 0.0;
            for (int c = 0; c < 3; c++) {
                r[c] = T.rc * T.chk[3 * i + c] - T.re * T.eq[3 * k + c];
                r2 += r[c] * r[c];
            }
            const double ri = 1.0 / std::sqrt(r2), ri3 = ri * ri * ri;
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++)
                    A[(size_t)(3 * k + b) * m + 3 * i + a] = ((a == b ? ri : 0.0) + r[a] * r[b] * ri3) / (8.0 * kPi);
        }
    // Householder QR: A = Q R; the reflectors v_j (v_j[j] = 1) overwrite A below the diagonal
    std::vector<double> tau(n), diag(n);
    for (int j = 0; j < n; j++) {
        double *a = &A[(size_t)j * m];
        double nrm = 0.0;
        for (int i = j; i < m; i++) nrm += a[i] * a[i];
        nrm = std::sqrt(nrm);
        SPHSTK_REQUIRE(nrm > 0.0, "internal: rank-deficient equivalent-source matrix");
        const double alpha = a[j] > 0 ? -nrm : nrm;
        const double v0 = a[j] - alpha;
        for (int i = j + 1; i < m; i++) a[i] /= v0;
        tau[j] = -v0 / alpha;
        diag[j] = alpha;
        a[j] = 1.0;
        for_columns(n - j - 1, [&](int jj) {
            double *b = &A[(size_t)(j + 1 + jj) * m];
            double s = 0.0;
            for (int i = j; i < m; i++) s += a[i] * b[i];
            s *= tau[j];
            for (int i = j; i < m; i++) b[i] -= s * a[i];
        });
    }
    // P = R^-1 Q^T: column c of Q^T is Q^T e_c (reflectors applied in order),
    // then back substitution with R (upper triangle of A, diagonal in diag)
    T.P.assign((size_t)n * m, 0.0);
    for_columns(m, [&](int c) {
        std::vector<double> y(m, 0.0);
        y[c] = 1.0;
        for (int j = 0; j < n; j++) {
            const double *a = &A[(size_t)j * m];
            double s = 0.0;
            for (int i = j; i < m; i++) s += a[i] * y[i];
            s *= tau[j];
            for (int i = j; i < m; i++) y[i] -= s * a[i];
        }
        for (int i = n - 1; i >= 0; i--) {
            double s = y[i];
            for (int k = i + 1; k < n; k++) s -= A[(size_t)k * m + i] * y[k];
            y[i] = s / diag[i];
        }
        for (int i = 0; i < n; i++) T.P[(size_t)i * m + c] = y[i];
    });
    return T;
}

} // namespace

const EquivTables &equiv_tables(int K) {
    SPHSTK_REQUIRE(K >= 16 && K <= 400, "equivalent-source template size out of range");
    // built once per size; entries of the map never move
    static std::mutex mu;
    static std::map<int, EquivTables> tabs;
    std::lock_guard<std::mutex> lk(mu);
    auto it = tabs.find(K);
    if (it == tabs.end()) it = tabs.emplace(K, build(K)).first;
    return it->second;
}

} // namespace sphstk
