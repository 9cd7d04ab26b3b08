// host_precompute.h - FP64 host-side precompute of libphpfdtd_hip.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/phpfdtd.h"

namespace phpfdtd {

constexpr double kEpsInfHost = 20.0;  // EPS_INF of the reference ADE (fdtd.c:202)

// CPML tables for one axis of n cells (fdtd.c:962-989 profile, per-index
// b/c/kappa of fdtd.c:3140-3216 hoisted out of the time loop).
struct AxisHost {
    std::vector<double> b, kh, kd, ch, cd;
    std::vector<int> slot;
};
AxisHost axis_tables(int n, double delta, const phpfdtd_grid &g);

// Coefficients of one material, all FP64.
struct MatPre {
    double R[9];     // lab -> crystal rotation (fdtd.c:3429-3452)
    double A[9];     // E_cr = A (D_cr - P)
    double M[9];     // E_lab = M D_lab (plain cells)
    double Lk[9];    // D loss matrix (fdtd.c:3349-3353)
    double Gpi[9], Gm[9], W[9];  // ADE recurrence
    double Ga[9], Gb[9];         // increment form: Gamma = gamma dt / 2, I - Gpi
    double inv_re[9], inv_im[9];
    double lf, lt;
    bool loss, special, ade;
    int natom;
};
int material_coefs(const phpfdtd_grid &g, const phpfdtd_material &m, MatPre *out, std::string *err);

int eps_inverse(const double a[9], const double b[9], double inv_re[9], double inv_im[9]);
int inverse3(const double m[9], double inv[9]);
double source_value(const phpfdtd_source &s, double t);
int validate_grid(const phpfdtd_grid &g, std::string *err);

}  // namespace phpfdtd
