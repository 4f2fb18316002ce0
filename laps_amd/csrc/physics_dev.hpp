// physics_dev.hpp -- device helpers shared by the physics and fused kernels.
#pragma once

#include "internal.hpp"

namespace laps {

// quotient in the arithmetic R of cfl_point: correctly rounded in FP64; the
// hardware reciprocal-based quotient (about 2 ulp) in FP32
template <typename R>
__device__ __forceinline__ R cfl_div(R a, R b) {
    if constexpr (sizeof(R) == 4) return __fdividef(a, b);
    else return a / b;
}

// square root in the arithmetic R of cfl_point: correctly rounded in FP64; the
// hardware square root (about 1 ulp) in FP32; x >= 0 here
template <typename R>
__device__ __forceinline__ R cfl_sqrt(R x) {
    if constexpr (sizeof(R) == 4) return __builtin_amdgcn_sqrtf(x);
    else return sqrt(x);
}

// dt candidate of one point (vardt, mhd.f90:355-409). R = double follows the
// reference's FP64 arithmetic; the mixed-precision step evaluates it in FP32
// (R = float), where the fields themselves are FP32. The reference takes per
// axis the largest of |u +- c_f|, |u +- c_s|, |u +- c_a| and |u|; with
// c_f >= c_a, c_s >= 0 (the fast speed bounds the others) that is |u| + c_f,
// which is what is computed here (equal up to the rounding of c_f against c_a
// in degenerate states, a few ulp of dt).
template <typename R>
__device__ __forceinline__ R cfl_point(R rho, R Bx, R By, R Bz, R p, R ux, R uy, R uz, const PhysParams& pp, R rr,
                                       R dmin) {
    const R cs2 = cfl_div(R(pp.gamma) * p, rho);
    const R sr = cfl_sqrt(rho);
    const R cax = cfl_div(Bx, sr), cay = cfl_div(By, sr), caz = cfl_div(Bz, sr);
    const R ca2 = cax * cax + cay * cay + caz * caz;
    const R cms2 = cs2 + ca2;
    const R zero = R(0);
    const R cnx = cfl_sqrt(fmax(cms2 * cms2 - 4 * cs2 * cax * cax, zero));
    const R cny = cfl_sqrt(fmax(cms2 * cms2 - 4 * cs2 * cay * cay, zero));
    const R cnz = cfl_sqrt(fmax(cms2 * cms2 - 4 * cs2 * caz * caz, zero));
    const R s2 = cfl_sqrt(R(2));
    const R cfx = cfl_div(cfl_sqrt(cms2 + cnx), s2), cfy = cfl_div(cfl_sqrt(cms2 + cny), s2);
    const R cfz = cfl_div(cfl_sqrt(cms2 + cnz), s2);
    R cmx = fabs(ux) + cfx, cmy = fabs(uy) + cfy, cmz = fabs(uz) + cfz;
    if (pp.hall) {
        R bmax = Bx;  // signed max of the components (mhd.f90:396-398)
        if (By > bmax) bmax = By;
        if (Bz > bmax) bmax = Bz;
        const R ch = cfl_div(cfl_div(R(pp.di), rho) * bmax, dmin);
        cmx = fmax(cmx, ch);
        cmy = fmax(cmy, ch);
        cmz = fmax(cmz, ch);
    }
    const R dtx = cfl_div(R(pp.dx), cmx), dty = cfl_div(R(pp.dy), cmy) * rr, dtz = cfl_div(R(pp.dz), cmz) * rr;
    R d = dtx < dty ? dtx : dty;
    return dtz < d ? dtz : d;
}

// k_square: grid_initialize form until the first evolve_radius, then
// update_ksquare (mhdinit.f90:114-122; AEBmod.f90:87-125); wave numbers w*
// r0r = R0/R, formed once per work-item (the reference's (k*R0)/R per mode, within an ulp)
__device__ __forceinline__ double ksquare(double kx, double ky, double kz, const PhysParams& pp, const Ctl* ctl,
                                          double r0r) {
    if (!ctl->ksq_updated) return kx * kx + ky * ky + kz * kz;
    if (pp.corot_ksq) {
        const double c = pp.cos_c, s = pp.sin_c;
        const double a = s * r0r, b = c * r0r, z = kz * r0r;
        return kx * kx * (c * c + a * a) + ky * ky * (s * s + b * b) + kx * ky * 2 * c * s * (1 - r0r * r0r) + z * z;
    }
    const double a = ky * r0r, b = kz * r0r;
    return kx * kx + a * a + b * b;
}

}  // namespace laps
