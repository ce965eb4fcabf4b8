/*
 * dev_resp.h -- device restatements of the reference's response taps:
 * z_resp_halfwidth (src/responses.c:27-46) and gen_z_response /
 * gen_r_response taps (src/responses.c:159-316), in FP64, rounded to FP32
 * as the reference stores them.  Shared by the refinement and the
 * local-power block planes.
 */
#ifndef DEV_RESP_H
#define DEV_RESP_H
#include "dev_fresnel.h"

/* z_resp_halfwidth, src/responses.c:27-46 */
__device__ static int zr_halfwidth(double z, int highacc)
{
    return (int) (0.5 * 1.1 * fabs(z)) + (highacc ? 48 : 16);
}

/* Tap m of gen_z_response(roffset, nb, z, numkern) (src/responses.c:228-316;
 * |z| < 1e-4: gen_r_response, :159-225), rounded to FP32. */
__device__ static float2 resp_tap_nb(double roffset, int nb, double z, int numkern, int m)
{
    const double PI = 3.14159265358979323846;
    double re, im;
    if (fabs(z) < 1e-4) {
        if (roffset < 1e-3 && m == numkern / 2) {
            double t = roffset * roffset;
            re = 1 - 6.579736267392905746 * t;
            im = roffset * (PI - 10.335425560099940058 * t);
        } else {
            double x = (numkern / 2 - m) / (double) nb + roffset;   /* response at pi x */
            double s, c;
            sincospi(x, &s, &c);
            double sinc = (x == 0.0) ? 1.0 : s / (PI * x);
            re = c * sinc;
            im = s * sinc;
        }
    } else {
        double ip;
        double startr = roffset - 0.5 * z;
        double sro = (startr < 0) ? 1.0 + modf(startr, &ip) : modf(startr, &ip);
        double absz = fabs(z);
        if (sro < 1e-3 && absz < 1e-3 && m == numkern / 2) {
            double zz = z * z, xx = sro * sro;
            re = 1.0 - 0.16449340668482264365 * zz;
            im = -0.5235987755982988731 * z;
            re += sro * 1.6449340668482264365 * z;
            im += sro * (PI - 0.5167712780049970029 * zz);
            re += xx * (-6.579736267392905746 + 0.9277056288952613070 * zz);
            im += xx * (3.1006276680299820175 * z);
        } else {
            int signz = (z < 0.0) ? -1 : 1;
            double zd = signz * 1.41421356237309504880 / sqrt(absz);
            double cons = zd / 2.0;
            double r = startr + (numkern / 2 - m) / (double) nb;
            double yy = r * zd, zz = yy + z * zd;
            double s, c;
            sincospi(r * r / z, &s, &c);
            double fsy, fcy, fsz, fcz;
            d_fresnl(yy, &fsy, &fcy);
            d_fresnl(zz, &fsz, &fcz);
            double tre = signz * (fcz - fcy), tim = fsy - fsz;
            re = (tre * c - tim * s) * cons;
            im = -(tre * s + tim * c) * cons;
        }
    }
    return make_float2((float) re, (float) im);
}

#endif
