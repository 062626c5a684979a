/* The C ABI from plain C99: include/sphstk.h compiles as C, the library links
 * and its device-independent entry points behave as declared.  Needs no GPU:
 * the collision LCP with the free-drain mobility is host code. */
#include "sphstk.h"

#include <math.h>
#include <stdio.h>
#include <string.h>

static int fails = 0;
#define CHECK(c)                                                    \
    do {                                                            \
        if (!(c)) {                                                 \
            fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #c); \
            fails++;                                                \
        }                                                           \
    } while (0)

int main(void) {
    CHECK(strncmp(sphstk_version(), "sphstk", 6) == 0);
    CHECK(sphstk_device_count() >= 0);
    /* errors come back as status codes with a message */
    int64_t n = 0;
    CHECK(sphstk_operator_sizes(NULL, &n, NULL, NULL, NULL) == SPHSTK_ERR_INVALID);
    CHECK(strlen(sphstk_last_error()) > 0);
    /* one overlapping pair, free drain: gamma = -b / A_11 in closed form */
    const int I[1] = {0}, J[1] = {1};
    const double phi0[1] = {-0.1}, nrm[3] = {-1.0, 0.0, 0.0}, radius[2] = {1.0, 0.5};
    double vk[12] = {0}, gamma[1] = {0}, F[12], U[12];
    sphstk_lcp_info info;
    const double dt = 0.01, mu = 0.9;
    CHECK(sphstk_collision_lcp(NULL, 2, radius, mu, 1, I, J, phi0, nrm, vk, dt, 0.0, 1e-12, 100, gamma, F, U,
                               &info) == SPHSTK_OK);
    const double pi = 3.14159265358979323846;
    const double a11 = dt * (1.0 / (6 * pi * mu * 1.0) + 1.0 / (6 * pi * mu * 0.5));
    CHECK(info.converged == 1);
    CHECK(fabs(gamma[0] - 0.1 / a11) < 1e-9 * (0.1 / a11));
    CHECK(fabs(F[0] + gamma[0]) < 1e-12 && fabs(F[6] - gamma[0]) < 1e-12);
    /* an invalid pair index is rejected before any work */
    const int Jbad[1] = {2};
    CHECK(sphstk_collision_lcp(NULL, 2, radius, mu, 1, I, Jbad, phi0, nrm, vk, dt, 0.0, 1e-12, 100, gamma, F, U,
                               &info) == SPHSTK_ERR_INVALID);
    if (fails) return 1;
    printf("ok\n");
    return 0;
}
