// host_precompute.cpp - FP64 host precompute of libphpfdtd_hip.
//
// Everything transcendental (exp of the CPML profile, cos/sin of the crystal
// rotations, the source waveforms) and every matrix inverse is evaluated here
// once, in FP64, so the device kernels contain no libm calls and the precision
// form only ever sees these values rounded once.
#include "host_precompute.h"

#include <math.h>
#include <string.h>

namespace phpfdtd {

static double det3(const double m[9]) {
    return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
           m[2] * (m[3] * m[7] - m[4] * m[6]);
}

// Adjugate inverse with the reference's singularity threshold |det| < 1e-12
// (fdtd.c:238-257).
int inverse3(const double m[9], double inv[9]) {
    const double det = det3(m);
    if (fabs(det) < 1e-12) return -1;
    const double c00 = m[4] * m[8] - m[5] * m[7], c01 = m[5] * m[6] - m[3] * m[8], c02 = m[3] * m[7] - m[4] * m[6];
    const double c10 = m[2] * m[7] - m[1] * m[8], c11 = m[0] * m[8] - m[2] * m[6], c12 = m[1] * m[6] - m[0] * m[7];
    const double c20 = m[1] * m[5] - m[2] * m[4], c21 = m[2] * m[3] - m[0] * m[5], c22 = m[0] * m[4] - m[1] * m[3];
    const double cof[9] = {c00, c01, c02, c10, c11, c12, c20, c21, c22};
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) inv[r * 3 + c] = cof[c * 3 + r] / det;
    return 0;
}

// The reference's "complex" inverse of eps = A + iB (fdtd.c:555-625): the
// cofactor matrices of A and B are combined with 1/(det A + i det B).  It is
// not the inverse of A + iB; only its real part enters the time step.
int eps_inverse(const double a[9], const double b[9], double inv_re[9], double inv_im[9]) {
    const double da = det3(a), db = det3(b);
    const double n2 = da * da + db * db;
    if (fabs(sqrt(n2)) < 1e-12) return -1;
    const double ir = da / n2, ii = -db / n2;
    auto cof_t = [](const double *m, double *o) {  // transposed cofactors (adjugate)
        o[0] = m[4] * m[8] - m[5] * m[7];
        o[1] = m[2] * m[7] - m[1] * m[8];
        o[2] = m[1] * m[5] - m[2] * m[4];
        o[3] = m[5] * m[6] - m[3] * m[8];
        o[4] = m[0] * m[8] - m[2] * m[6];
        o[5] = m[2] * m[3] - m[0] * m[5];
        o[6] = m[3] * m[7] - m[4] * m[6];
        o[7] = m[1] * m[6] - m[0] * m[7];
        o[8] = m[0] * m[4] - m[1] * m[3];
    };
    double ca[9], cb[9];
    cof_t(a, ca);
    cof_t(b, cb);
    for (int e = 0; e < 9; ++e) {
        inv_re[e] = ir * ca[e] - ii * cb[e];
        inv_im[e] = ir * cb[e] + ii * ca[e];
    }
    return 0;
}

static void matmul3(const double *a, const double *b, double *c) {
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k) c[r * 3 + k] = a[r * 3] * b[k] + a[r * 3 + 1] * b[3 + k] + a[r * 3 + 2] * b[6 + k];
}

static void transpose3(const double *a, double *t) {
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) t[c * 3 + r] = a[r * 3 + c];
}

AxisHost axis_tables(int n, double delta, const phpfdtd_grid &g) {
    AxisHost ax;
    const int L = g.pml_thickness;
    ax.b.assign(n, 0.0);
    ax.kh.assign(n, 0.0);
    ax.kd.assign(n, 0.0);
    ax.ch.assign(n, 0.0);
    ax.cd.assign(n, 0.0);
    ax.slot.assign(n, -1);
    // graded profile, fdtd.c:971-986
    std::vector<double> sig(L), kap(L), alp(L);
    if (L > 0) {
        const double eta0 = sqrt(g.mu0 / g.eps0);
        const double smax = -(g.pml_order + 1) * log(g.pml_reflection) / (2.0 * eta0 * delta * L);
        for (int p = 0; p < L; ++p) {
            const double r = (double)(L - p) / L;
            const double rm = pow(r, g.pml_order);
            sig[p] = smax * rm;
            kap[p] = 1.0 + (g.kappa_max - 1.0) * rm;
            alp[p] = g.alpha_max * (1.0 - r);
        }
    }
    for (int i = 0; i < n; ++i) {
        int p = -1, slot = -1;
        if (L > 0 && i < L) {  // low side wins on overlap, as the reference's if/else-if
            p = i;
            slot = i;
        } else if (L > 0 && i >= n - L) {
            p = n - 1 - i;
            slot = L + (i - (n - L));  // ascending in i, so a lane's x vector is contiguous in its slab
        }
        double kappa = 1.0, b = 0.0, c = 0.0;
        if (p >= 0) {
            kappa = kap[p];
            const double aw = sig[p] / (kappa * g.eps0) + alp[p] / g.eps0;
            b = exp(-aw * g.dt);
            c = (aw > 1e-30) ? sig[p] / (kappa * kappa * g.eps0) * (b - 1.0) / aw : 0.0;
        }
        ax.slot[i] = slot;
        ax.b[i] = b;
        ax.kh[i] = g.dt / (g.mu0 * delta * kappa);
        ax.kd[i] = g.dt / (delta * kappa);
        ax.ch[i] = g.dt * c / (g.mu0 * delta);
        ax.cd[i] = g.dt * c / delta;
    }
    return ax;
}

int material_coefs(const phpfdtd_grid &g, const phpfdtd_material &m, MatPre *o, std::string *err) {
    memset(o, 0, sizeof(*o));
    if (eps_inverse(m.eps_re, m.eps_im, o->inv_re, o->inv_im) != 0) {
        *err = "Error: Singular dielectric tensor!";
        return PHPFDTD_EINVAL;
    }
    if (fabs(m.anglex) > 1e-10 && fabs(m.anglez) > 1e-10) {
        *err = "Error: Can not rotate along X and Z axis at the same time";
        return PHPFDTD_EINVAL;
    }
    if (m.use_ade != 0 && m.use_ade != 1) {
        *err = "Error: use_ADE must be 0 or 1";
        return PHPFDTD_EINVAL;
    }
    if (m.natom < 0 || m.natom > 32) {
        *err = "natom must be in [0, 32]";
        return PHPFDTD_EINVAL;
    }
    if (m.natom > 0 && (!m.qezv || !m.qezm || !m.dmm)) {
        *err = "Error: QeZV, QeZM and DMM must be set when natom > 0";
        return PHPFDTD_EINVAL;
    }
    o->natom = m.natom;
    o->ade = m.use_ade == 1;
    o->special = o->ade || m.natom > 0;
    // conduction loss factors, fdtd.c:3413-3415
    const double x = m.sigma * g.dt / (2 * g.eps0);
    o->lf = 1.0 / (1.0 + x);
    o->lt = x * o->lf;
    // rotation into the crystal frame, fdtd.c:3429-3452 (anglez takes precedence)
    double R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    if (fabs(m.anglez) > 1e-10) {
        const double c = cos(m.anglez), s = sin(m.anglez);
        const double r[9] = {c, -s, 0, s, c, 0, 0, 0, 1};
        memcpy(R, r, sizeof(R));
    } else if (fabs(m.anglex) > 1e-10) {
        const double c = cos(-m.anglex), s = sin(-m.anglex);
        const double r[9] = {1, 0, 0, 0, c, -s, 0, s, c};
        memcpy(R, r, sizeof(R));
    }
    memcpy(o->R, R, sizeof(R));
    if (o->ade) {
        // Drude fit of fdtd.c:2478-2528 at w = 2 pi f_source
        const double w = g.f_source * 2.0 * M_PI;
        double e1[9], e1i[9], gam[9], gp[9];
        memcpy(e1, m.eps_re, sizeof(e1));
        e1[0] -= kEpsInfHost;
        e1[4] -= kEpsInfHost;
        e1[8] -= kEpsInfHost;
        if (e1[0] >= 0 || e1[4] >= 0 || e1[8] >= 0) {
            *err = "Error: EPS_INF is too small";
            return PHPFDTD_EINVAL;
        }
        if (inverse3(e1, e1i) != 0) {
            *err = "Error: epsr1 can not be inversed";
            return PHPFDTD_EINVAL;
        }
        double ei_e1i[9];
        matmul3(m.eps_im, e1i, ei_e1i);
        for (int e = 0; e < 9; ++e) gam[e] = -ei_e1i[e] * w;
        double gei[9];
        matmul3(gam, m.eps_im, gei);
        double wp2[9];
        for (int e = 0; e < 9; ++e) wp2[e] = -w * w * e1[e] + w * gei[e];
        for (int e = 0; e < 9; ++e) {
            gp[e] = gam[e] * g.dt / 2.0 + ((e % 4) == 0 ? 1.0 : 0.0);
            o->Gm[e] = -gam[e] * g.dt / 2.0 + ((e % 4) == 0 ? 1.0 : 0.0);
            o->W[e] = g.eps0 * g.dt * g.dt * wp2[e];
            o->Ga[e] = gam[e] * g.dt / 2.0;
        }
        if (inverse3(gp, o->Gpi) != 0) {
            *err = "Error: gdt2p1 can not be inversed";
            return PHPFDTD_EINVAL;
        }
        // I - Gpi = Gamma (I + Gamma)^-1, the small matrix of the increment form
        for (int e = 0; e < 9; ++e) o->Gb[e] = ((e % 4) == 0 ? 1.0 : 0.0) - o->Gpi[e];
        const double f = (1.0 - o->lt) / (g.eps0 * kEpsInfHost);
        for (int e = 0; e < 9; ++e) o->A[e] = ((e % 4) == 0) ? f : 0.0;
    } else {
        const double f = (1.0 - o->lt) * o->lf / g.eps0;
        for (int e = 0; e < 9; ++e) o->A[e] = f * o->inv_re[e];
        bool any = false;
        for (int e = 0; e < 9; ++e) any |= m.eps_im[e] != 0.0;
        o->loss = any;
        // fdtd.c:3350 uses the literal 3.14159265
        const double wl = g.dt * g.f_source * 2.0 * 3.14159265 * g.eps0;
        for (int e = 0; e < 9; ++e) o->Lk[e] = wl * m.eps_im[e];
    }
    // plain cells: E_lab = R^T A R D_lab
    double Rt[9], AR[9];
    transpose3(R, Rt);
    matmul3(o->A, R, AR);
    matmul3(Rt, AR, o->M);
    return 0;
}

// fdtd.c:1923-1947
double source_value(const phpfdtd_source &s, double t) {
    if (t < s.start_time || t > s.end_time) return 0.0;
    switch (s.type) {
        case 0: {
            const double u = (t - s.t0) / s.tau;
            return s.amplitude * exp(-(u * u));
        }
        case 1:
            return s.amplitude * sin(2 * M_PI * s.frequency * t);
        case 2: {
            const double ts = t - s.t0;
            const double arg = M_PI * M_PI * s.frequency * s.frequency * ts * ts;
            return s.amplitude * (1.0 - 2.0 * arg) * exp(-arg);
        }
        case 3: {
            const double u = (t - s.t0) / s.tau;
            return s.amplitude * sin(2 * M_PI * s.frequency * t) * exp(-(u * u));
        }
        default:
            return 0.0;
    }
}

int validate_grid(const phpfdtd_grid &g, std::string *err) {
    if (g.nx < 3 || g.ny < 3 || g.nz < 1) {
        *err = "grid must be at least 3 x 3 x 1 cells";
        return PHPFDTD_EINVAL;
    }
    // slot cell indices are uint32 offsets into the padded field arrays (rows
    // of nx rounded up to 32 cells), so the padded count must fit
    if ((int64_t)((g.nx + 31) / 32 * 32) * g.ny * g.nz > (int64_t)UINT32_MAX) {
        *err = "grid exceeds 2^32 cells (rows padded to a multiple of 32 cells)";
        return PHPFDTD_EINVAL;
    }
    if (!(g.dx > 0 && g.dy > 0 && g.dz > 0 && g.dt > 0 && g.eps0 > 0 && g.mu0 > 0)) {
        *err = "dx, dy, dz, dt, eps0 and mu0 must be positive";
        return PHPFDTD_EINVAL;
    }
    if (g.pml_thickness < 0 || g.tsteps < 0) {
        *err = "pml_thickness and tsteps must be non-negative";
        return PHPFDTD_EINVAL;
    }
    if (g.pml_thickness > 0 && !(g.pml_reflection > 0 && g.pml_reflection < 1)) {
        *err = "pml_reflection must be in (0, 1)";
        return PHPFDTD_EINVAL;
    }
    if (g.precision != PHPFDTD_FP64 && g.precision != PHPFDTD_FP32 && g.precision != PHPFDTD_FP32_RV64) {
        *err = "precision must be PHPFDTD_FP64, PHPFDTD_FP32 or PHPFDTD_FP32_RV64";
        return PHPFDTD_EINVAL;
    }
    return 0;
}

}  // namespace phpfdtd
