/*
 * Test sources for the tree builder, with the calling convention of the
 * reference's `fun` (every argument by reference): the same compiled
 * function is handed to this library's builder and to the reference's, so
 * both see bit-identical values at every node.
 *
 * dpars: ng rows of (c[ndim], s[ndim], a); ipars: ndim, ng, iperiod.
 * Density q of nd: sum_g a exp(-sum_k (x_k - c_k)^2 / s_k) * (1 + q cos(3 x_1 + q)),
 * summed over the 3^ndim nearest lattice images when iperiod = 1.
 */
#include <math.h>

void test_source(int *nd, double *xyz, double *dpars, void *zpars, int *ipars, double *f)
{
    (void)zpars;
    const int ndim = ipars[0], ng = ipars[1], iper = ipars[2];
    const int stride = 2 * ndim + 1;
    int nimg = 1;
    for (int k = 0; k < ndim; ++k) nimg *= iper ? 3 : 1;
    double acc = 0.0;
    for (int g = 0; g < ng; ++g) {
        const double *p = dpars + g * stride;
        for (int im = 0; im < nimg; ++im) {
            int code = im;
            double rr = 0.0;
            for (int k = 0; k < ndim; ++k) {
                double shift = 0.0;
                if (iper) { shift = (double)(code % 3 - 1); code /= 3; }
                const double d = xyz[k] - p[k] - shift;
                rr += d * d / p[ndim + k];
            }
            acc += p[2 * ndim] * exp(-rr);
        }
    }
    for (int q = 0; q < *nd; ++q) f[q] = acc * (1.0 + q * cos(3.0 * xyz[0] + q));
}
