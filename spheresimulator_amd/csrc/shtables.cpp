// Host-side spherical-harmonic set-up: Gauss-Legendre rule, the per-order unit
// grid of Sphere/Shexp.cpp:206-271 (poles dropped as in
// BIE/SphereSTKSHOperator.cpp:103-121) and the analysis matrix that maps
// body-frame grid values to vector-SH coefficients (the role of SCTL
// Grid2VecSHC behind Shexp::calcSpectralCoeff, Sphere/Shexp.cpp:408-447).
//
// Basis (per degree n<=p, order 0<=m<=n; m<0 implied by conjugation):
//   y_nm = D_n^m on the unit sphere, D_n^m = (d/dx + i d/dy)^m (d/dz)^(n-m) (1/r)
//   e_r = y r,  e_s = grad_s y,  e_t = r x grad_s y,
// with |y_nm|^2 integrating to 4 pi (n-m)!(n+m)!/(2n+1) and |e_s|^2, |e_t|^2 to
// n(n+1) times that.  The analysis is the orthogonal projection in the grid's
// quadrature inner product, which is exact for vector-SH degree <= p.
#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <complex>
#include <map>
#include <memory>
#include <mutex>

namespace sphstk {

using cplx = std::complex<double>;

void gauss_legendre(int n, std::vector<double> &x, std::vector<double> &w) {
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    for (int i = 0; i < (n + 1) / 2; i++) {
        // Newton iteration on P_n starting from the Tricomi approximation of root i (descending)
        double z = std::cos(kPi * (4.0 * i + 3.0) / (4.0 * n + 2.0));
        double pn = 0, dpn = 1;
        for (int it = 0; it < 60; it++) {
            double pm1 = 1.0, p0 = z;
            if (n == 0) p0 = 1.0;
            for (int k = 1; k < n; k++) {
                const double pk1 = ((2 * k + 1) * z * p0 - k * pm1) / (k + 1);
                pm1 = p0;
                p0 = pk1;
            }
            pn = p0;
            dpn = n * (pm1 - z * pn) / (1.0 - z * z);
            const double dz = pn / dpn;
            z -= dz;
            if (std::fabs(dz) <= 1e-17) break;
        }
        double pm1 = 1.0, p0 = z;
        for (int k = 1; k < n; k++) {
            const double pk1 = ((2 * k + 1) * z * p0 - k * pm1) / (k + 1);
            pm1 = p0;
            p0 = pk1;
        }
        dpn = n * (pm1 - z * p0) / (1.0 - z * z);
        x[n - 1 - i] = z;
        x[i] = -z;
        w[i] = w[n - 1 - i] = 2.0 / ((1.0 - z * z) * dpn * dpn);
    }
    if (n % 2) x[n / 2] = 0.0;
}

// D_n^m(x), 0<=m<=n<=P, by the diagonal and the three-term z recurrence.
static void solid_harmonics(int P, const double *x, std::vector<cplx> &D) {
    D.assign(tri_size(P), cplx(0, 0));
    const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
    const double ir2 = 1.0 / r2;
    const cplx xy(x[0], x[1]);
    D[0] = 1.0 / std::sqrt(r2);
    for (int m = 1; m <= P; m++) D[tri(m, m)] = -(2.0 * m - 1.0) * ir2 * xy * D[tri(m - 1, m - 1)];
    for (int m = 0; m < P; m++) {
        cplx prev = 0.0, cur = D[tri(m, m)];
        for (int n = m; n < P; n++) {
            const cplx next = -((2.0 * n + 1.0) * x[2] * cur + double(n * n - m * m) * prev) * ir2;
            D[tri(n + 1, m)] = next;
            prev = cur;
            cur = next;
        }
    }
}

// Rows of the synthesis operator at one unit vector e (body frame): the value
// component c of the expansion at e is sum_j row[c * 6Q + j] coef[j].
static void synthesis_rows_at(int p, const double *e, std::vector<cplx> &D, double *row) {
    const int Q = tri_size(p), ncoef = 6 * Q;
    std::fill(row, row + 3 * ncoef, 0.0);
    solid_harmonics(p + 1, e, D);
    for (int n = 0; n <= p; n++)
        for (int m = 0; m <= n; m++) {
            const int k = tri(n, m);
            cplx g[3];
            g[2] = D[tri(n + 1, m)];
            if (m > 0) {
                g[0] = 0.5 * (D[tri(n + 1, m + 1)] - D[tri(n + 1, m - 1)]);
                g[1] = cplx(0, -0.5) * (D[tri(n + 1, m + 1)] + D[tri(n + 1, m - 1)]);
            } else {
                g[0] = D[tri(n + 1, 1)].real();
                g[1] = D[tri(n + 1, 1)].imag();
            }
            const cplx y = D[k];
            cplx gs[3], xt[3];
            for (int c = 0; c < 3; c++) gs[c] = g[c] + (n + 1.0) * y * e[c];
            xt[0] = e[1] * g[2] - e[2] * g[1];
            xt[1] = e[2] * g[0] - e[0] * g[2];
            xt[2] = e[0] * g[1] - e[1] * g[0];
            const double eps = m ? 2.0 : 1.0;
            for (int c = 0; c < 3; c++) {
                double *rc = row + size_t(c) * ncoef;
                const cplx br = y * e[c], bs = gs[c], bt = xt[c];
                rc[2 * k] = eps * br.real();
                rc[2 * k + 1] = -eps * br.imag();
                rc[2 * Q + 2 * k] = eps * bs.real();
                rc[2 * Q + 2 * k + 1] = -eps * bs.imag();
                rc[4 * Q + 2 * k] = eps * bt.real();
                rc[4 * Q + 2 * k + 1] = -eps * bt.imag();
            }
        }
}

std::vector<double> pole_rows(int p) {
    const int ncoef = 6 * tri_size(p);
    std::vector<double> rows(size_t(6) * ncoef);
    std::vector<cplx> D;
    const double north[3] = {0.0, 0.0, 1.0}, south[3] = {0.0, 0.0, -1.0};
    synthesis_rows_at(p, north, D, rows.data());
    synthesis_rows_at(p, south, D, rows.data() + size_t(3) * ncoef);
    return rows;
}

static std::unique_ptr<ShTables> build_tables(int p) {
    auto t = std::make_unique<ShTables>();
    t->p = p;
    t->npts = grid_points(p);
    t->Q = tri_size(p);
    std::vector<double> xg, wg;
    gauss_legendre(p + 1, xg, wg);
    const int nphi = 2 * p + 2;
    t->eta.resize(3 * t->npts);
    t->w.resize(t->npts);
    for (int j = 0; j <= p; j++) {
        const double ct = xg[p - j], st = std::sqrt(1.0 - ct * ct);
        for (int k = 0; k < nphi; k++) {
            const double ph = 2.0 * kPi * k / nphi;
            const int q = j * nphi + k;
            t->eta[3 * q] = st * std::cos(ph);
            t->eta[3 * q + 1] = st * std::sin(ph);
            t->eta[3 * q + 2] = ct;
            t->w[q] = 2.0 * kPi / nphi * wg[p - j];
        }
    }
    const int Q = t->Q, rows = 6 * Q, cols = 3 * t->npts;
    t->anal.assign(size_t(rows) * cols, 0.0);
    std::vector<double> invN(Q);
    for (int n = 0; n <= p; n++)
        for (int m = 0; m <= n; m++) {
            double f = 4.0 * kPi / (2.0 * n + 1.0);
            for (int k = 2; k <= n - m; k++) f *= k;
            for (int k = 2; k <= n + m; k++) f *= k;
            invN[tri(n, m)] = 1.0 / f;
        }
    t->synth.assign(size_t(cols) * rows, 0.0);
    std::vector<cplx> D;
    for (int q = 0; q < t->npts; q++) {
        const double *e = &t->eta[3 * q];
        solid_harmonics(p + 1, e, D);
        for (int n = 0; n <= p; n++)
            for (int m = 0; m <= n; m++) {
                const int k = tri(n, m);
                // Cartesian gradient of D_n^m from degree n+1
                cplx g[3];
                g[2] = D[tri(n + 1, m)];
                if (m > 0) {
                    g[0] = 0.5 * (D[tri(n + 1, m + 1)] - D[tri(n + 1, m - 1)]);
                    g[1] = cplx(0, -0.5) * (D[tri(n + 1, m + 1)] + D[tri(n + 1, m - 1)]);
                } else {
                    g[0] = D[tri(n + 1, 1)].real();
                    g[1] = D[tri(n + 1, 1)].imag();
                }
                const cplx y = D[k];
                cplx gs[3], xt[3];
                for (int c = 0; c < 3; c++) gs[c] = g[c] + (n + 1.0) * y * e[c];
                xt[0] = e[1] * g[2] - e[2] * g[1];
                xt[1] = e[2] * g[0] - e[0] * g[2];
                xt[2] = e[0] * g[1] - e[1] * g[0];
                // synthesis: value_c = sum_k eps_m Re(F_r y e_c + F_s gs_c + F_t xt_c), F stored (re, im)
                const double eps = m ? 2.0 : 1.0;
                for (int c = 0; c < 3; c++) {
                    const size_t r = size_t(3 * q + c);
                    const cplx br = y * e[c], bs = gs[c], bt = xt[c];
                    t->synth[size_t(2 * k) * cols + r] = eps * br.real();
                    t->synth[size_t(2 * k + 1) * cols + r] = -eps * br.imag();
                    t->synth[size_t(2 * Q + 2 * k) * cols + r] = eps * bs.real();
                    t->synth[size_t(2 * Q + 2 * k + 1) * cols + r] = -eps * bs.imag();
                    t->synth[size_t(4 * Q + 2 * k) * cols + r] = eps * bt.real();
                    t->synth[size_t(4 * Q + 2 * k + 1) * cols + r] = -eps * bt.imag();
                }
                const double wr = t->w[q] * invN[k];
                const double wt = n > 0 ? t->w[q] * invN[k] / (n * (n + 1.0)) : 0.0;
                for (int c = 0; c < 3; c++) {
                    const size_t col = size_t(3 * q + c) * rows;
                    const cplx cr = std::conj(y) * e[c] * wr;
                    const cplx cs = std::conj(gs[c]) * wt;
                    const cplx ctt = std::conj(xt[c]) * wt;
                    t->anal[col + 2 * k] = cr.real();
                    t->anal[col + 2 * Q + 2 * k] = cs.real();
                    t->anal[col + 4 * Q + 2 * k] = ctt.real();
                    if (m > 0) {
                        t->anal[col + 2 * k + 1] = cr.imag();
                        t->anal[col + 2 * Q + 2 * k + 1] = cs.imag();
                        t->anal[col + 4 * Q + 2 * k + 1] = ctt.imag();
                    }
                }
            }
    }
    return t;
}

// rows of a coefficient vector (6Q layout) that belong to degrees <= pc
static std::vector<int> coef_rows_upto(int p, int pc) {
    const int Q = tri_size(p);
    std::vector<int> r;
    for (int blk = 0; blk < 3; blk++)
        for (int n = 0; n <= std::min(p, pc); n++)
            for (int m = 0; m <= n; m++) {
                r.push_back(blk * 2 * Q + 2 * tri(n, m));
                r.push_back(blk * 2 * Q + 2 * tri(n, m) + 1);
            }
    return r;
}

static std::unique_ptr<Resample> build_resample(int p, int pc) {
    auto rs = std::make_unique<Resample>();
    rs->p = p;
    rs->pc = pc;
    const ShTables &F = sh_tables(p), &C = sh_tables(rs->pc);
    rs->n = F.npts;
    rs->nc = C.npts;
    const int QF = F.Q, QC = C.Q;
    const int rowsF = 6 * QF, rowsC = 6 * QC;
    if (rs->pc == p) { // same grid: identity in both directions
        const int nn = 3 * F.npts;
        rs->down.assign(size_t(nn) * nn, 0.0);
        for (int i = 0; i < nn; i++) rs->down[size_t(i) * nn + i] = 1.0;
        rs->up = rs->down;
        return rs;
    }
    const std::vector<int> rf = coef_rows_upto(p, rs->pc), rc = coef_rows_upto(rs->pc, rs->pc);
    // down = synth_C[:, rc] * anal_F[rf, :]
    const int nF = 3 * F.npts, nC = 3 * C.npts;
    rs->down.assign(size_t(nC) * nF, 0.0);
    for (int col = 0; col < nF; col++)
        for (size_t k = 0; k < rf.size(); k++) {
            const double a = F.anal[size_t(col) * rowsF + rf[k]];
            if (a == 0.0) continue;
            const double *sc = &C.synth[size_t(rc[k]) * nC];
            double *o = &rs->down[size_t(col) * nC];
            for (int r = 0; r < nC; r++) o[r] += sc[r] * a;
        }
    // up = synth_F[:, rf] * anal_C[rc, :]
    rs->up.assign(size_t(nF) * nC, 0.0);
    for (int col = 0; col < nC; col++)
        for (size_t k = 0; k < rc.size(); k++) {
            const double a = C.anal[size_t(col) * rowsC + rc[k]];
            if (a == 0.0) continue;
            const double *sf = &F.synth[size_t(rf[k]) * nF];
            double *o = &rs->up[size_t(col) * nF];
            for (int r = 0; r < nF; r++) o[r] += sf[r] * a;
        }
    return rs;
}

const Resample &resample_tables(int p, int pc) {
    static std::mutex mu;
    static std::map<std::pair<int, int>, std::unique_ptr<Resample>> cache;
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find({p, pc});
    if (it != cache.end()) return *it->second;
    auto t = build_resample(p, pc);
    const Resample &ref = *t;
    cache[{p, pc}] = std::move(t);
    return ref;
}

const ShTables &sh_tables(int p) {
    static std::mutex mu;
    static std::map<int, std::unique_ptr<ShTables>> cache;
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(p);
    if (it != cache.end()) return *it->second;
    auto t = build_tables(p);
    const ShTables &ref = *t;
    cache[p] = std::move(t);
    return ref;
}

} // namespace sphstk
