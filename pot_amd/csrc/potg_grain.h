// potg_grain.h -- analytic field and potential of the spheroidal grain
// (pot-spheroid.c), shared by the walks and the direct sum.
#pragma once

#include "potg_dev.h"

namespace potg {

// Field of the charged conducting spheroid, axis along x
// (field_particle_feels_from_grain, SPH:1689-1780).
__device__ inline void grain_field(const InstParams& p, int32_t gq, double x, double y, double z,
                            double* E)
{
    double c, phi_const, rt, E_n, beta;
    const double A = p.A;
    if (p.geometry == 1) {
        c = (p.a * sqrt(A * A - 1.0)) * pow(A, -1.0 / 3.0);
        phi_const = (QE * gq * acosh(A)) / (4.0 * PI * EPSILON0 * c * log(A - sqrt(A * A - 1.0)));
    } else {
        c = (p.a * sqrt(1.0 - A * A)) * pow(A, -1.0 / 3.0);
        phi_const = (QE * gq * acos(A)) /
                    (8.0 * PI * EPSILON0 * c * (atan((1.0 - sqrt(1.0 - A * A)) / A) - 0.25 * PI));
    }
    const double dx = x - p.grain[0], dy = y - p.grain[1], dz = z - p.grain[2];
    const double rho2 = (dy * dy) + (dz * dz);
    if (p.geometry == 1) {
        rt = sqrt(rho2 * rho2 + dx * dx * dx * dx + c * c * c * c + 2.0 * rho2 * dx * dx +
                  2.0 * rho2 * c * c - 2.0 * dx * dx * c * c);
        const double sinh2 = (rho2 + dx * dx - c * c + rt) / (2.0 * c * c);
        E_n = -phi_const / sqrt(sinh2 * rt);
        beta = atan((-sinh2 * dx) / ((sinh2 + 1.0) * sqrt(rho2)));
    } else {
        rt = sqrt(rho2 * rho2 + dx * dx * dx * dx + c * c * c * c + 2.0 * rho2 * dx * dx -
                  2.0 * rho2 * c * c + 2.0 * dx * dx * c * c);
        const double cosh2 = (rho2 + dx * dx + c * c + rt) / (2.0 * c * c);
        E_n = -phi_const / sqrt(cosh2 * rt);
        beta = atan((-cosh2 * dx) / ((cosh2 - 1.0) * sqrt(rho2)));
    }
    const double theta = atan(fabs(dz) / fabs(dy));
    E[0] = -E_n * sin(beta);
    E[1] = E_n * cos(beta) * cos(theta) * fabs(dy) / dy;
    E[2] = E_n * cos(beta) * sin(theta) * fabs(dz) / dz;
}

// Spheroid potential at absolute coordinates, as phi_at seeds it (SPH:1924-1953).
__device__ inline double grain_phi_abs(const InstParams& p, int32_t gq, double x, double y, double z)
{
    const double A = p.A;
    const double rho2 = (y * y) + (z * z);
    if (p.geometry == 1) {
        const double c = (p.a * sqrt(A * A - 1.0)) / pow(A, 1.0 / 3.0);
        const double phi_const =
            (QE * gq * acosh(A)) / (4.0 * PI * EPSILON0 * c * log(A - sqrt(A * A - 1.0)));
        const double rt = sqrt(rho2 * rho2 + x * x * x * x + c * c * c * c + 2.0 * rho2 * x * x +
                               2.0 * rho2 * c * c - 2.0 * x * x * c * c);
        return phi_const *
               log((sqrt(rho2 + x * x + c * c + rt) - c * sqrt(2.0)) / sqrt(rho2 + x * x - c * c + rt));
    }
    const double c = (p.a * sqrt(1.0 - A * A)) * pow(A, -1.0 / 3.0);
    const double phi_const =
        (QE * gq * acos(A)) /
        (4.0 * PI * EPSILON0 * c * (atan((1.0 - sqrt(1.0 - A * A)) / A) - 0.25 * PI));
    const double rt = sqrt(rho2 * rho2 + x * x * x * x + c * c * c * c + 2.0 * rho2 * x * x -
                           2.0 * rho2 * c * c + 2.0 * x * x * c * c);
    return phi_const * (atan((sqrt(rho2 + x * x + c * c + rt) - c * sqrt(2.0)) /
                             sqrt(rho2 + x * x - c * c + rt)) -
                        0.25 * PI);
}

}  // namespace potg
