/*
 * Closed-form per-bin integrals of the 2-D static blob family in FP64: the truth the
 * RMSE of an FP32 result is measured against, where the integrand has one.
 *   Gaussian  int exp(-a (x-c)^2) dx = sqrt(pi/a)/2 [erf(sqrt(a)(x1-c)) - erf(sqrt(a)(x0-c))]
 *             (erfc differences on the tail side keep the digits)
 *   wave      A int sin(fx x) dx int cos(fy y) dy, separable
 *   step      h * area of {nx x + ny y > d} inside the bin: the bin polygon clipped
 *             by the half-plane (Sutherland-Hodgman), then the shoelace formula
 * A bin's value is (bin count) * (integral over the bin), the scaling of the
 * integrators' bins (factor = number of bins in
 * src/control-variates/regions-integrator-parallel-variance-reduction.h:77-88).
 */
#ifndef VMX_EXACT_H
#define VMX_EXACT_H

#include <hip/hip_runtime.h>

#include "region_math.h"

namespace vmx {

__device__ inline double gauss_segment(double x0, double x1, double c, double a) {
    if (a <= 0.0) return x1 - x0;
    const double r = sqrt(a);
    const double u0 = r * (x0 - c), u1 = r * (x1 - c);
    double d;
    if (u0 >= 0.0)
        d = erfc(u0) - erfc(u1);
    else if (u1 <= 0.0)
        d = erfc(-u1) - erfc(-u0);
    else
        d = erf(u1) - erf(u0);
    return 0.5 * sqrt(3.14159265358979323846 / a) * d;
}

/* area of {nx x + ny y > d} within [x0,x1] x [y0,y1] */
__device__ inline double halfplane_area(double x0, double x1, double y0, double y1, double nx, double ny, double d) {
    double px[8], py[8], qx[8], qy[8];
    px[0] = x0; py[0] = y0;
    px[1] = x1; py[1] = y0;
    px[2] = x1; py[2] = y1;
    px[3] = x0; py[3] = y1;
    int n = 4, m = 0;
    for (int i = 0; i < n; ++i) {
        const int j = (i + 1) % n;
        const double si = nx * px[i] + ny * py[i] - d, sj = nx * px[j] + ny * py[j] - d;
        const bool ini = si > 0.0, inj = sj > 0.0;
        if (ini) {
            qx[m] = px[i];
            qy[m] = py[i];
            ++m;
        }
        if (ini != inj) {
            const double t = si / (si - sj);
            qx[m] = px[i] + t * (px[j] - px[i]);
            qy[m] = py[i] + t * (py[j] - py[i]);
            ++m;
        }
    }
    double a2 = 0.0;
    for (int i = 0; i < m; ++i) {
        const int j = (i + 1) % m;
        a2 += qx[i] * qy[j] - qx[j] * qy[i];
    }
    return 0.5 * fabs(a2);
}

struct ExactBlobs2D {
    int nblobs, step_mode;
    double blob[VMX_MAX_BLOBS][7];
    double step[5], wave[3];
};

static __global__ void k_exact_bins(const ExactBlobs2D* __restrict__ p, double lo0, double lo1, double dx, double dy,
                             uint64_t res0, uint64_t res1, double* __restrict__ out) {
    const uint64_t b = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (b >= res0 * res1) return;
    const uint64_t i = b / res1, j = b % res1;
    const double x0 = lo0 + double(i) * dx, x1 = lo0 + double(i + 1) * dx;
    const double y0 = lo1 + double(j) * dy, y1 = lo1 + double(j + 1) * dy;
    double s = 0.0;
    for (int k = 0; k < p->nblobs; ++k) {
        const double* q = p->blob[k];
        s += q[6] * gauss_segment(x0, x1, q[0], q[4]) * gauss_segment(y0, y1, q[1], q[5]);
    }
    const double A = p->wave[0], fx = p->wave[1], fy = p->wave[2];
    if (A != 0.0) {
        const double ix = fx == 0.0 ? 0.0 : (cos(fx * x0) - cos(fx * x1)) / fx;
        const double iy = fy == 0.0 ? (y1 - y0) : (sin(fy * y1) - sin(fy * y0)) / fy;
        s += A * ix * iy;
    }
    if (p->step_mode == 1) s += p->step[4] * halfplane_area(x0, x1, y0, y1, p->step[0], p->step[1], p->step[3]);
    out[b] = double(res0 * res1) * s;
}

}  // namespace vmx

#endif
