/* Runs the FP64 oracle (oracle/ctprj_oracle.c) under -fsanitize=address,undefined:
 * forward, adjoint, FBP, nnz and the dense matrix on small parallel and fan geometries
 * (even and odd sizes, partial arcs, pitch != 1), with 1 and 3 threads, and checks
 * <Ax, y> = <x, A^T y> and that the dense matrix reproduces the forward.  Exit status 0
 * when every check holds. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "ctprj_oracle.h"

static int failures = 0;

static double frand(unsigned* s)
{
    *s = *s * 1103515245u + 12345u;
    return ((*s >> 8) & 0xFFFF) / 65536.0 - 0.5;
}

static void run(orc_conf c, int nth)
{
    const int ni = c.n * c.n, ns = c.np * c.prjWidth;
    double* x = malloc(sizeof(double) * ni);
    double* y = malloc(sizeof(double) * ns);
    double* Ax = malloc(sizeof(double) * ns);
    double* Aty = malloc(sizeof(double) * ni);
    double* fb = malloc(sizeof(double) * ni);
    double* A = malloc(sizeof(double) * (size_t)ni * ns);
    unsigned s = 7u + (unsigned)c.n;
    int i, r;
    for (i = 0; i < ni; i++) x[i] = frand(&s);
    for (i = 0; i < ns; i++) y[i] = frand(&s);
    const long long nf = orc_forward(&c, x, Ax, nth);
    const long long na = orc_adjoint(&c, y, Aty, nth);
    const long long nz = orc_nnz(&c, nth);
    orc_fbp(&c, y, fb, nth);
    orc_dense(&c, A);
    double lhs = 0, rhs = 0, mag = 0, err = 0;
    for (i = 0; i < ns; i++) lhs += Ax[i] * y[i];
    for (i = 0; i < ni; i++) rhs += x[i] * Aty[i];
    for (r = 0; r < ns; r++) {
        double v = 0;
        for (i = 0; i < ni; i++) v += A[(size_t)r * ni + i] * x[i];
        err = fmax(err, fabs(v - Ax[r]));
        mag = fmax(mag, fabs(Ax[r]));
    }
    if (!(nf == na && na == nz && nz > 0) || fabs(lhs - rhs) > 1e-12 * (fabs(lhs) + 1) ||
        err > 1e-12 * (mag + 1)) {
        fprintf(stderr, "n=%d W=%d np=%d P=%d d=%g: nnz %lld/%lld/%lld, <Ax,y>-<x,Aty> %g, "
                "dense %g\n", c.n, c.prjWidth, c.np, c.prjFull, c.d, nf, na, nz, lhs - rhs, err);
        failures++;
    }
    free(x);
    free(y);
    free(Ax);
    free(Aty);
    free(fb);
    free(A);
}

int main(void)
{
    const orc_conf cases[] = {
        {8, 13, 12, 24, 1.0, 1.0, 0.0, 1, 0},
        {9, 12, 10, 36, 1.0, 0.8, 40.0, 1, 0},
        {7, 11, 5, 20, 1.3, 1.0, 0.0, 2, 1},
        {10, 17, 9, 9, 0.9, 1.0, 30.0, 4, 0},
    };
    int k;
    for (k = 0; k < (int)(sizeof(cases) / sizeof(cases[0])); k++) {
        run(cases[k], 1);
        run(cases[k], 3);
    }
    if (failures) fprintf(stderr, "%d cases failed\n", failures);
    else printf("oracle driver: all checks passed\n");
    return failures ? 1 : 0;
}
