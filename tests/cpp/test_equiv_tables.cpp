// Host-side check of the equivalent-source templates of the far-field cell
// aggregation (equiv.cpp): P is a left inverse of the unit cell's Stokeslet
// matrix, and the fitted point forces reproduce the velocity and the traction
// of 16 clustered sphere grids (radius 1/8 inside the unit cell, random
// forces), with the total force (the flow's leading Stokeslet term) preserved.
// Argument K: 200 (the FP32 operator's template, targets at 4 cell radii, to
// the accuracy that operator needs) or 32 (the coarse operator's, targets on
// its check sphere at 2.5 cell radii, to a few 1e-3 in velocity: white-noise
// forces are its worst case).
#include "common.hpp"

#include <cmath>
#include <cstdio>
#include <array>
#include <random>
#include <cstdlib>

using namespace sphstk;

static void stokeslet(const double *r, const double *f, double *u) {
    const double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2], ri = 1.0 / std::sqrt(r2);
    const double rf = r[0] * f[0] + r[1] * f[1] + r[2] * f[2];
    for (int a = 0; a < 3; a++) u[a] += (f[a] * ri + r[a] * rf * ri * ri * ri) / (8 * kPi);
}
static void traction(const double *r, const double *n, const double *f, double *t) {
    const double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2], ri = 1.0 / std::sqrt(r2);
    const double rf = r[0] * f[0] + r[1] * f[1] + r[2] * f[2], rn = r[0] * n[0] + r[1] * n[1] + r[2] * n[2];
    for (int a = 0; a < 3; a++) t[a] += -3.0 / (4 * kPi) * r[a] * rf * rn * std::pow(ri, 5);
}

int main(int argc, char **argv) {
    const int K = argc > 1 ? std::atoi(argv[1]) : kEquivK;
    const bool coarse = K == 32;
    if (K != kEquivK && !coarse) return 2;
    const double dist = coarse ? 2.5 : 4.0;
    const EquivTables &T = equiv_tables(K);
    const int n = 3 * T.K, m = 3 * T.Kc;
    int fails = 0;
    // P G = I
    double eI = 0.0;
    for (int k = 0; k < T.K; k++)
        for (int b = 0; b < 3; b++) {
            std::vector<double> col(m, 0.0);
            const double fb[3] = {b == 0 ? 1.0 : 0.0, b == 1 ? 1.0 : 0.0, b == 2 ? 1.0 : 0.0};
            for (int i = 0; i < T.Kc; i++) {
                double r[3];
                for (int c = 0; c < 3; c++) r[c] = T.rc * T.chk[3 * i + c] - T.re * T.eq[3 * k + c];
                stokeslet(r, fb, &col[3 * i]);
            }
            for (int row = 0; row < n; row++) {
                double s = 0.0;
                for (int j = 0; j < m; j++) s += T.P[(size_t)row * m + j] * col[j];
                eI = std::fmax(eI, std::fabs(s - (row == 3 * k + b ? 1.0 : 0.0)));
            }
        }
    std::printf("max |P G - I| = %.2e\n", eI);
    fails += !(eI < 1e-6);
    // a clustered cell
    std::mt19937_64 rng(7);
    std::normal_distribution<double> N;
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    const ShTables &g12 = sh_tables(12);
    std::vector<double> y, f;
    std::vector<std::array<double, 3>> cen;
    while (cen.size() < 16) {
        std::array<double, 3> c = {0.875 * U(rng), 0.875 * U(rng), 0.875 * U(rng)};
        if (std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]) > 0.875) continue;
        bool ok = true;
        for (auto &d : cen) ok = ok && std::hypot(c[0] - d[0], std::hypot(c[1] - d[1], c[2] - d[2])) > 0.25;
        if (!ok) continue;
        cen.push_back(c);
        for (int q = 0; q < g12.npts; q++)
            for (int a = 0; a < 3; a++) {
                y.push_back(c[a] + 0.125 * g12.eta[3 * q + a]);
                f.push_back(N(rng));
            }
    }
    const int ns = (int)y.size() / 3;
    std::vector<double> uc(m, 0.0), g(n, 0.0);
    for (int i = 0; i < T.Kc; i++)
        for (int s = 0; s < ns; s++) {
            double r[3];
            for (int c = 0; c < 3; c++) r[c] = T.rc * T.chk[3 * i + c] - y[3 * s + c];
            stokeslet(r, &f[3 * s], &uc[3 * i]);
        }
    for (int row = 0; row < n; row++)
        for (int j = 0; j < m; j++) g[row] += T.P[(size_t)row * m + j] * uc[j];
    double Fs[3] = {0, 0, 0}, Fg[3] = {0, 0, 0};
    for (int s = 0; s < ns; s++)
        for (int c = 0; c < 3; c++) Fs[c] += f[3 * s + c];
    for (int k = 0; k < T.K; k++)
        for (int c = 0; c < 3; c++) Fg[c] += g[3 * k + c];
    const double eF = std::hypot(Fs[0] - Fg[0], std::hypot(Fs[1] - Fg[1], Fs[2] - Fg[2])) /
                      std::hypot(Fs[0], std::hypot(Fs[1], Fs[2]));
    double nu = 0, du = 0, nt = 0, dt = 0;
    for (int it = 0; it < 300; it++) {
        double x[3] = {N(rng), N(rng), N(rng)}, nn[3] = {N(rng), N(rng), N(rng)};
        const double lx = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
        const double ln = std::sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
        for (int c = 0; c < 3; c++) {
            x[c] *= dist / lx;
            nn[c] /= ln;
        }
        double ue[3] = {0, 0, 0}, ud[3] = {0, 0, 0}, te[3] = {0, 0, 0}, td[3] = {0, 0, 0};
        for (int k = 0; k < T.K; k++) {
            double r[3];
            for (int c = 0; c < 3; c++) r[c] = x[c] - T.re * T.eq[3 * k + c];
            stokeslet(r, &g[3 * k], ue);
            traction(r, nn, &g[3 * k], te);
        }
        for (int s = 0; s < ns; s++) {
            double r[3];
            for (int c = 0; c < 3; c++) r[c] = x[c] - y[3 * s + c];
            stokeslet(r, &f[3 * s], ud);
            traction(r, nn, &f[3 * s], td);
        }
        for (int c = 0; c < 3; c++) {
            nu += ud[c] * ud[c];
            du += (ue[c] - ud[c]) * (ue[c] - ud[c]);
            nt += td[c] * td[c];
            dt += (te[c] - td[c]) * (te[c] - td[c]);
        }
    }
    const double eu = std::sqrt(du / nu), et = std::sqrt(dt / nt);
    std::printf("K = %d, 16 clustered sphere grids, targets at %.1f rho: velocity %.2e traction %.2e total force "
                "%.2e\n", K, dist, eu, et, eF);
    if (coarse)
        fails += !(eu < 5e-3 && et < 4e-2 && eF < 2e-4);
    else
        fails += !(eu < 1e-6 && et < 5e-6 && eF < 1e-8);
    return fails ? 1 : 0;
}
