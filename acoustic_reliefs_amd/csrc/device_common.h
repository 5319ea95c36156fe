// Device helpers shared by the gfx950 kernels: complex arithmetic on double2/float2, the
// 6-point Green's-function quadrature entry, wavefront reductions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "trig.h"

#define ARH_WAVE 64

namespace arh {

// Degree-4, 6-point triangle rule: weight, barycentric coordinates of b and c
// (constants.h:39-46; the table must equal the reference's value for value).
#define ARH_QW_A 0.223381589678011
#define ARH_QW_B 0.109951743655322
#define ARH_Q_1 0.445948490915965
#define ARH_Q_2 0.108103018168070
#define ARH_Q_3 0.091576213509771
#define ARH_Q_4 0.816847572980459

__device__ __forceinline__ void quad_point(int g, double &w, double &xi, double &eta) {
    switch (g) {
    case 0: w = ARH_QW_A; xi = ARH_Q_1; eta = ARH_Q_2; break;
    case 1: w = ARH_QW_A; xi = ARH_Q_1; eta = ARH_Q_1; break;
    case 2: w = ARH_QW_A; xi = ARH_Q_2; eta = ARH_Q_1; break;
    case 3: w = ARH_QW_B; xi = ARH_Q_3; eta = ARH_Q_4; break;
    case 4: w = ARH_QW_B; xi = ARH_Q_3; eta = ARH_Q_3; break;
    default: w = ARH_QW_B; xi = ARH_Q_4; eta = ARH_Q_3; break;
    }
}

constexpr double kPi = 3.14159265358979323846;
constexpr double kInv8Pi = 1.0 / (8.0 * kPi);  // (1/2 rule factor) * 1/(4 pi)

// Per-element geometry record (16 doubles = 128 B): a, b - a, c - a, n = (b-a)x(c-a), centroid.
struct Tri {
    double ax, ay, az, bx, by, bz, cx, cy, cz, nx, ny, nz, gx, gy, gz, pad;
};
static_assert(sizeof(Tri) == 128, "Tri layout");

// M(p, triangle) = 1/2 sum_g w_g K(p, q_g, n),  K = e^{-ikr}(-ikr-1)/(4 pi r^2) * (-(p-q).n / r)
// (integrate.h:11-19, integrate.cpp:5-14). With d = p - q, s = -(d.n):
// e^{-ikr}(-1-ikr) = (-cos - kr sin) + i(sin - kr cos).
__device__ __forceinline__ double2 bem_entry(double k, double px, double py, double pz, const Tri &t) {
    double re = 0.0, im = 0.0;
#pragma unroll
    for (int g = 0; g < 6; g++) {
        double w, xi, eta;
        quad_point(g, w, xi, eta);
        double qx = t.ax + xi * t.bx + eta * t.cx;
        double qy = t.ay + xi * t.by + eta * t.cy;
        double qz = t.az + xi * t.bz + eta * t.cz;
        double dx = px - qx, dy = py - qy, dz = pz - qz;
        double r2 = dx * dx + dy * dy + dz * dz;
        double r = sqrt(r2);
        double s = -(dx * t.nx + dy * t.ny + dz * t.nz);
        double kr = k * r;
        double sn, cs;
        trig_sincos(kr, &sn, &cs);
        double f = w * s / (r2 * r);
        re += f * (-cs - kr * sn);
        im += f * (sn - kr * cs);
    }
    return make_double2(re * kInv8Pi, im * kInv8Pi);
}

__device__ __forceinline__ Tri load_tri(const Tri *__restrict__ tri, int e) { return tri[e]; }

// ---------------------------------------------------------------- complex helpers
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// acc += a * b
__device__ __forceinline__ void cfma(double2 &acc, double2 a, double2 b) {
    acc.x = fma(a.x, b.x, acc.x);
    acc.x = fma(-a.y, b.y, acc.x);
    acc.y = fma(a.x, b.y, acc.y);
    acc.y = fma(a.y, b.x, acc.y);
}
// conj(a) * b
__device__ __forceinline__ double2 cconjmul(double2 a, double2 b) {
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double2 cdiv(double2 a, double2 b) {
    // Smith's algorithm
    if (fabs(b.x) >= fabs(b.y)) {
        double r = b.y / b.x, den = b.x + b.y * r;
        return make_double2((a.x + a.y * r) / den, (a.y - a.x * r) / den);
    } else {
        double r = b.x / b.y, den = b.y + b.x * r;
        return make_double2((a.x * r + a.y) / den, (a.y * r - a.x) / den);
    }
}
__device__ __forceinline__ double cnorm2(double2 a) { return a.x * a.x + a.y * a.y; }
__device__ __forceinline__ double cabs_(double2 a) { return hypot(a.x, a.y); }

// ---------------------------------------------------------------- wavefront reductions (64 lanes)
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double2 wave_sum2(double2 v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        v.x += __shfl_xor(v.x, o, 64);
        v.y += __shfl_xor(v.y, o, 64);
    }
    return v;
}
// (value, index) max with ties to the lower index; value -1 = empty.
__device__ __forceinline__ void wave_argmax(double &v, int &i) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        double ov = __shfl_xor(v, o, 64);
        int oi = __shfl_xor(i, o, 64);
        if (ov > v || (ov == v && oi < i)) {
            v = ov;
            i = oi;
        }
    }
}

}  // namespace arh
