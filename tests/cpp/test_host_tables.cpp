// Host-side check of the product's spherical-harmonic tables (shtables.cpp):
// the resampling matrices between an order-p grid and an order-pc grid are
// exact on band-limited fields (constant, rigid rotation, gradient of x^3) and
// remove degrees above pc; the analysis/synthesis pair is a projection.
#include "common.hpp"

#include <cmath>
#include <cstdio>

using namespace sphstk;

static void field(const double *e, double *v, int kind) {
    if (kind == 0) {
        v[0] = 0.3, v[1] = -1.1, v[2] = 0.7;
    } else if (kind == 1) {
        const double c[3] = {0.4, 0.1, -0.9};
        v[0] = c[1] * e[2] - c[2] * e[1], v[1] = c[2] * e[0] - c[0] * e[2], v[2] = c[0] * e[1] - c[1] * e[0];
    } else {
        v[0] = 3 * e[0] * e[0], v[1] = e[1] * e[2], v[2] = 0;
    }
}

int main() {
    int fails = 0;
    for (int p : {6, 8, 12, 16})
        for (int pc : {4, 5, p / 2 + 1, p}) {
            const Resample &rs = resample_tables(p, pc);
            const ShTables &F = sh_tables(p), &C = sh_tables(pc);
            double errd = 0, erru = 0;
            for (int kind = 0; kind < 3; kind++) {
                std::vector<double> f(3 * rs.n), c(3 * rs.nc), d(3 * rs.nc, 0), u(3 * rs.n, 0);
                for (int q = 0; q < rs.n; q++) field(&F.eta[3 * q], &f[3 * q], kind);
                for (int q = 0; q < rs.nc; q++) field(&C.eta[3 * q], &c[3 * q], kind);
                for (int col = 0; col < 3 * rs.n; col++)
                    for (int r = 0; r < 3 * rs.nc; r++) d[r] += rs.down[(size_t)col * 3 * rs.nc + r] * f[col];
                for (int col = 0; col < 3 * rs.nc; col++)
                    for (int r = 0; r < 3 * rs.n; r++) u[r] += rs.up[(size_t)col * 3 * rs.n + r] * c[col];
                for (int r = 0; r < 3 * rs.nc; r++) errd = std::fmax(errd, std::fabs(d[r] - c[r]));
                for (int r = 0; r < 3 * rs.n; r++) erru = std::fmax(erru, std::fabs(u[r] - f[r]));
            }
            // a degree-p radial field y_p0 r is removed entirely when pc < p
            std::vector<double> f(3 * rs.n), d(3 * rs.nc, 0);
            for (int q = 0; q < rs.n; q++) {
                const double z = F.eta[3 * q + 2];
                double P0 = 1, P1 = z;
                for (int k = 1; k < p; k++) {
                    const double P2 = ((2 * k + 1) * z * P1 - k * P0) / (k + 1);
                    P0 = P1;
                    P1 = P2;
                }
                for (int c = 0; c < 3; c++) f[3 * q + c] = P1 * F.eta[3 * q + c];
            }
            for (int col = 0; col < 3 * rs.n; col++)
                for (int r = 0; r < 3 * rs.nc; r++) d[r] += rs.down[(size_t)col * 3 * rs.nc + r] * f[col];
            double tr = 0;
            for (double v : d) tr = std::fmax(tr, std::fabs(v));
            const bool ok = errd < 1e-13 && erru < 1e-13 && (pc == p || tr < 1e-13);
            std::printf("p=%d pc=%d: down %.2e up %.2e truncation %.2e %s\n", p, pc, errd, erru, tr, ok ? "ok" : "FAIL");
            fails += !ok;
        }
    return fails ? 1 : 0;
}
