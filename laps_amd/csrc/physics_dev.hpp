// physics_dev.hpp -- device helpers shared by the physics and fused kernels.
#pragma once

#include "internal.hpp"

namespace laps {

// dt candidate of one point (vardt, mhd.f90:355-409), FP64
__device__ __forceinline__ double cfl_point(double rho, double Bx, double By, double Bz, double p, double ux,
                                            double uy, double uz, const PhysParams& pp, double rr, double dmin) {
    const double cs2 = pp.gamma * p / rho;
    const double sr = sqrt(rho);
    const double cax = Bx / sr, cay = By / sr, caz = Bz / sr;
    const double ca2 = cax * cax + cay * cay + caz * caz;
    const double cms2 = cs2 + ca2;
    const double cnx = sqrt(fmax(cms2 * cms2 - 4 * cs2 * cax * cax, 0.0));
    const double cny = sqrt(fmax(cms2 * cms2 - 4 * cs2 * cay * cay, 0.0));
    const double cnz = sqrt(fmax(cms2 * cms2 - 4 * cs2 * caz * caz, 0.0));
    const double s2 = sqrt(2.0);
    const double cfx = sqrt(cms2 + cnx) / s2, cfy = sqrt(cms2 + cny) / s2, cfz = sqrt(cms2 + cnz) / s2;
    const double csx = sqrt(fmax(cms2 - cnx, 0.0)) / s2;
    const double csy = sqrt(fmax(cms2 - cny, 0.0)) / s2;
    const double csz = sqrt(fmax(cms2 - cnz, 0.0)) / s2;
    auto cmax = [](double u, double cf, double cs, double ca) {
        return fmax(fmax(fmax(fabs(u + cf), fabs(u + cs)), fmax(fabs(u + ca), fabs(u - cf))),
                    fmax(fmax(fabs(u - cs), fabs(u - ca)), fabs(u)));
    };
    double cmx = cmax(ux, cfx, csx, cax), cmy = cmax(uy, cfy, csy, cay), cmz = cmax(uz, cfz, csz, caz);
    if (pp.hall) {
        double bmax = Bx;  // signed max of the components (mhd.f90:396-398)
        if (By > bmax) bmax = By;
        if (Bz > bmax) bmax = Bz;
        const double ch = pp.di / rho * bmax / dmin;
        cmx = fmax(cmx, ch);
        cmy = fmax(cmy, ch);
        cmz = fmax(cmz, ch);
    }
    const double dtx = pp.dx / cmx, dty = pp.dy / cmy * rr, dtz = pp.dz / cmz * rr;
    double d = dtx < dty ? dtx : dty;
    return dtz < d ? dtz : d;
}

// k_square: grid_initialize form until the first evolve_radius, then
// update_ksquare (mhdinit.f90:114-122; AEBmod.f90:87-125); wave numbers w*
__device__ __forceinline__ double ksquare(double kx, double ky, double kz, const PhysParams& pp, const Ctl* ctl,
                                          double radius) {
    if (!ctl->ksq_updated) return kx * kx + ky * ky + kz * kz;
    const double R0 = pp.radius0, R = radius;
    if (pp.corot_ksq) {
        const double c = pp.cos_c, s = pp.sin_c, r0r = R0 / R;
        const double a = s * R0 / R, b = c * R0 / R, z = kz * R0 / R;
        return kx * kx * (c * c + a * a) + ky * ky * (s * s + b * b) + kx * ky * 2 * c * s * (1 - r0r * r0r) + z * z;
    }
    const double a = ky * R0 / R, b = kz * R0 / R;
    return kx * kx + a * a + b * b;
}

}  // namespace laps
