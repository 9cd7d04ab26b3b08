/*
 * fdtd_oracle.c - CPU FP64 restatement of the PhP_FDTD time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product (php_fdtd_amd, libphpfdtd_hip) never links, loads or falls back
 * to it.
 *
 * It restates, from the mathematics and the reference's step ordering, the
 * single-rank behaviour of fdtd.c (reference file, cited by line):
 *   init_pml                      fdtd.c:962-989
 *   inverse / determinant         fdtd.c:205-258
 *   inverse_dielectric_tensor     fdtd.c:555-625   (pseudo-complex)
 *   ADE coefficients              fdtd.c:2478-2528
 *   calculate_source_value        fdtd.c:1923-1947
 *   H update                      fdtd.c:3123-3231 (lagged psi)
 *   D update + eps_im loss        fdtd.c:3238-3356 (unrotated eps_im)
 *   source injection              fdtd.c:3363-3396
 *   D->E, ADE, ion EOM            fdtd.c:3404-3531
 *   diagnostics                   fdtd.c:3535-3596
 *   monitor sampling (z+1)        fdtd.c:3806-3851
 * with one deliberate difference: the ADE update uses the intended full 3x3
 * recurrence (all of P^{n-1} and of the right-hand side before the solve).
 * The reference's loop (fdtd.c:3465-3475) reads uninitialised tempvec entries
 * and partially updated P_ADEo; it is defined, and equal to this one bit for
 * bit, only when gamma and wp2 are diagonal.
 *
 * How each function is pinned (each check named is a test that exists):
 *   tests/test_oracle.py, against the reference binary built from fdtd.c by
 *   oracle/Makefile (golden runs under tests/golden/): every monitor row,
 *   full-monitor row (all r, v) and float32 slice bit for bit, over CPML
 *   with unequal cells, anisotropic lossy rotated media, diagonal ADE from
 *   keys and from a switchxz table, the hBN and MoO3 lattices and the
 *   built-in materials without CPML.  This pins init_pml, the pseudo-complex
 *   eps^-1, the ADE coefficients and recurrence for diagonal tensors, the
 *   sources, the H/D/E updates, the ion EOM and the sampling planes.
 *   tests/test_oracle_pins.py, for what those runs leave open: orc_diag
 *   against numpy and against the values the reference printed (including
 *   its divergence report), orc_inverse3 by A A^-1 = I, the ADE
 *   coefficients for full tensors by the Drude identity, the full-tensor
 *   ADE recurrence against numpy's linear solve on one cell, and the ion
 *   EOM against the eigenmode recurrences of a non-symmetric DMM.
 *
 * Layout: the reference's own, z fastest with a one-plane zero halo on each z
 * side: (i,j,k) -> (i*ny + j)*(nz+2) + k + 1  (fdtd.c:261-284).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/phpfdtd.h"

#define ORC_EPS_INF 20.0 /* fdtd.c:202 */

/* ------------------------------------------------------------------------ */
/* 3x3 algebra (fdtd.c:205-258)                                             */
/* ------------------------------------------------------------------------ */
double orc_det3(const double m[9]) {
    return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
           m[2] * (m[3] * m[7] - m[4] * m[6]);
}

/* inv = adj^T / det; returns -1 when |det| < 1e-12 (fdtd.c:238-257). */
int orc_inverse3(const double m[9], double inv[9]) {
    double det = orc_det3(m);
    if (fabs(det) < 1e-12) return -1;
    double adj[9];
    adj[0] = (m[4] * m[8] - m[5] * m[7]);
    adj[1] = -(m[3] * m[8] - m[5] * m[6]);
    adj[2] = (m[3] * m[7] - m[4] * m[6]);
    adj[3] = -(m[1] * m[8] - m[2] * m[7]);
    adj[4] = (m[0] * m[8] - m[2] * m[6]);
    adj[5] = -(m[0] * m[7] - m[1] * m[6]);
    adj[6] = (m[1] * m[5] - m[2] * m[4]);
    adj[7] = -(m[0] * m[5] - m[2] * m[3]);
    adj[8] = (m[0] * m[4] - m[1] * m[3]);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) inv[i * 3 + j] = adj[j * 3 + i] / det;
    return 0;
}

/* Pseudo-complex inverse of eps = A + iB (fdtd.c:555-625).  Not the true
 * complex inverse: det(A) and det(B) are combined as if they were the real and
 * imaginary parts of det(A+iB), and cofactors of A and B are mixed per entry.
 * Only the real part is used by the time step (fdtd.c:3459-3461).
 * Returns -1 for the reference's "Singular dielectric tensor" exit. */
int orc_eps_inverse(const double a[9], const double b[9], double inv_re[9], double inv_im[9]) {
    double dr = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                a[2] * (a[3] * a[7] - a[4] * a[6]);
    double di = b[0] * (b[4] * b[8] - b[5] * b[7]) - b[1] * (b[3] * b[8] - b[5] * b[6]) +
                b[2] * (b[3] * b[7] - b[4] * b[6]);
    double mag = sqrt(dr * dr + di * di);
    if (fabs(mag) < 1e-12) return -1;
    double ir = dr / (dr * dr + di * di);
    double ii = -di / (dr * dr + di * di);
    /* cofactors in the entry pattern of fdtd.c:581-622 */
    double ca[9], cb[9];
    const double *M;
    M = a;
    ca[0] = M[4] * M[8] - M[5] * M[7];
    ca[1] = M[2] * M[7] - M[1] * M[8];
    ca[2] = M[1] * M[5] - M[2] * M[4];
    ca[3] = M[5] * M[6] - M[3] * M[8];
    ca[4] = M[0] * M[8] - M[2] * M[6];
    ca[5] = M[2] * M[3] - M[0] * M[5];
    ca[6] = M[3] * M[7] - M[4] * M[6];
    ca[7] = M[1] * M[6] - M[0] * M[7];
    ca[8] = M[0] * M[4] - M[1] * M[3];
    M = b;
    cb[0] = M[4] * M[8] - M[5] * M[7];
    cb[1] = M[2] * M[7] - M[1] * M[8];
    cb[2] = M[1] * M[5] - M[2] * M[4];
    cb[3] = M[5] * M[6] - M[3] * M[8];
    cb[4] = M[0] * M[8] - M[2] * M[6];
    cb[5] = M[2] * M[3] - M[0] * M[5];
    cb[6] = M[3] * M[7] - M[4] * M[6];
    cb[7] = M[1] * M[6] - M[0] * M[7];
    cb[8] = M[0] * M[4] - M[1] * M[3];
    for (int e = 0; e < 9; e++) {
        inv_re[e] = ir * ca[e] - ii * cb[e];
        inv_im[e] = ir * cb[e] + ii * ca[e];
    }
    return 0;
}

/* CPML profile (fdtd.c:962-989): sigma_max = -(m+1) ln R / (2 eta0 delta L),
 * r = (L-i)/L, sigma = sigma_max r^m, kappa = 1 + (kmax-1) r^m,
 * alpha = amax (1-r). */
void orc_pml_profile(int L, double delta, double eps0, double mu0, double refl, double m,
                     double kmax, double amax, double *sigma, double *kappa, double *alpha) {
    double eta0 = sqrt(mu0 / eps0);
    double smax = -(m + 1) * log(refl) / (2.0 * eta0 * delta * L);
    for (int i = 0; i < L; i++) {
        double r = (double)(L - i) / L;
        sigma[i] = smax * pow(r, m);
        kappa[i] = 1.0 + (kmax - 1.0) * pow(r, m);
        alpha[i] = amax * (1.0 - r);
    }
}

/* Source waveform (fdtd.c:1923-1947). */
double orc_source_value(const phpfdtd_source *s, double time) {
    if (time < s->start_time || time > s->end_time) return 0.0;
    switch (s->type) {
    case 0:
        return s->amplitude * exp(-pow((time - s->t0) / s->tau, 2));
    case 1:
        return s->amplitude * sin(2 * M_PI * s->frequency * time);
    case 2: {
        double ts = time - s->t0;
        double arg = M_PI * M_PI * s->frequency * s->frequency * ts * ts;
        return s->amplitude * (1.0 - 2.0 * arg) * exp(-arg);
    }
    case 3:
        return s->amplitude * sin(2 * M_PI * s->frequency * time) *
               exp(-pow((time - s->t0) / s->tau, 2));
    default:
        return 0.0;
    }
}

/* ADE coefficients (fdtd.c:2478-2528), w = 2 pi f:
 *   epsr1 = eps_re - EPS_INF I;  gamma = -w eps_im epsr1^-1;
 *   wp2 = -w^2 epsr1 + w gamma eps_im;  G+- = I +- gamma dt/2.
 * Returns -1 for "EPS_INF is too small", -2 for singular epsr1. */
int orc_ade_coefs(const double er[9], const double ei[9], double f, double dt, double gama[9],
                  double wp2[9], double gp[9], double gm[9], double gp_inv[9], double gm_inv[9]) {
    double e1[9], e1i[9];
    memcpy(e1, er, sizeof(e1));
    e1[0] -= ORC_EPS_INF;
    e1[4] -= ORC_EPS_INF;
    e1[8] -= ORC_EPS_INF;
    if (e1[0] >= 0 || e1[4] >= 0 || e1[8] >= 0) return -1;
    if (orc_inverse3(e1, e1i) != 0) return -2;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++)
            gama[r * 3 + c] = -(ei[r * 3 + 0] * e1i[0 * 3 + c] + ei[r * 3 + 1] * e1i[1 * 3 + c] +
                                ei[r * 3 + 2] * e1i[2 * 3 + c]) *
                              f * 2.0 * M_PI;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++)
            wp2[r * 3 + c] = -f * f * 4.0 * M_PI * M_PI * e1[r * 3 + c] +
                             f * 2.0 * M_PI *
                                 (gama[r * 3 + 0] * ei[0 * 3 + c] + gama[r * 3 + 1] * ei[1 * 3 + c] +
                                  gama[r * 3 + 2] * ei[2 * 3 + c]);
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            gp[r * 3 + c] = +gama[r * 3 + c] * dt / 2.0;
            gm[r * 3 + c] = -gama[r * 3 + c] * dt / 2.0;
            if (r == c) {
                gp[r * 3 + c] += 1.0;
                gm[r * 3 + c] += 1.0;
            }
        }
    /* the reference only prints when these are singular (fdtd.c:2518-2527) */
    if (orc_inverse3(gm, gm_inv) != 0) memset(gm_inv, 0, 9 * sizeof(double));
    if (orc_inverse3(gp, gp_inv) != 0) memset(gp_inv, 0, 9 * sizeof(double));
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Simulation state                                                         */
/* ------------------------------------------------------------------------ */
typedef struct {
    phpfdtd_material m;
    double *qezv, *qezm, *dmm;
    double inv_re[9], inv_im[9];
    double wp2[9], gm[9], gp_inv[9];
} orc_mat;

typedef struct orc_sim {
    phpfdtd_grid g;
    int nx, ny, nz, L;
    size_t nfield;
    double *F[9]; /* Ex Ey Ez Dx Dy Dz Hx Hy Hz */
    double *psiHxy, *psiHxz, *psiDxy, *psiDxz; /* [2L][ny][nz] */
    double *psiHyx, *psiHyz, *psiDyx, *psiDyz; /* [nx][2L][nz] */
    double *psiHzx, *psiHzy, *psiDzx, *psiDzy; /* [nx][ny][nz] */
    double *pml_s[3], *pml_k[3], *pml_a[3];
    int nmat;
    orc_mat *mats;
    int32_t *medium; /* [nx][ny][nz] */
    int64_t *ion_off; /* per cell offset into r/v (-1 none) */
    int64_t *ade_off; /* per cell offset into P/Po (-1 none) */
    double *r, *v, *P, *Po;
    int nsrc;
    phpfdtd_source *src;
    long step;
} orc_sim;

#define FI(s, i, j, k) ((((size_t)(i)) * (s)->ny + (j)) * ((s)->nz + 2) + (k) + 1)
#define CI(s, i, j, k) ((((size_t)(i)) * (s)->ny + (j)) * (s)->nz + (k))

static void set_err(char *err, int n, const char *msg) {
    if (err && n > 0) snprintf(err, n, "%s", msg);
}

void orc_destroy(orc_sim *s) {
    if (!s) return;
    for (int c = 0; c < 9; c++) free(s->F[c]);
    free(s->psiHxy); free(s->psiHxz); free(s->psiDxy); free(s->psiDxz);
    free(s->psiHyx); free(s->psiHyz); free(s->psiDyx); free(s->psiDyz);
    free(s->psiHzx); free(s->psiHzy); free(s->psiDzx); free(s->psiDzy);
    for (int d = 0; d < 3; d++) { free(s->pml_s[d]); free(s->pml_k[d]); free(s->pml_a[d]); }
    if (s->mats)
        for (int m = 0; m < s->nmat; m++) {
            free(s->mats[m].qezv); free(s->mats[m].qezm); free(s->mats[m].dmm);
        }
    free(s->mats); free(s->medium); free(s->ion_off); free(s->ade_off);
    free(s->r); free(s->v); free(s->P); free(s->Po); free(s->src);
    free(s);
}

orc_sim *orc_create(const phpfdtd_grid *g, const phpfdtd_material *mats, int nmat,
                    const int32_t *medium, char *err, int errlen) {
    orc_sim *s = (orc_sim *)calloc(1, sizeof(orc_sim));
    if (!s) return NULL;
    s->g = *g;
    s->nx = g->nx; s->ny = g->ny; s->nz = g->nz; s->L = g->pml_thickness;
    int nx = s->nx, ny = s->ny, nz = s->nz, L = s->L;
    size_t ncell = (size_t)nx * ny * nz;
    s->nfield = (size_t)nx * ny * (nz + 2);
    for (int c = 0; c < 9; c++) s->F[c] = (double *)calloc(s->nfield, sizeof(double));
    size_t nxs = (size_t)(2 * L) * ny * nz, nys = (size_t)nx * (2 * L) * nz;
    s->psiHxy = calloc(nxs + 1, 8); s->psiHxz = calloc(nxs + 1, 8);
    s->psiDxy = calloc(nxs + 1, 8); s->psiDxz = calloc(nxs + 1, 8);
    s->psiHyx = calloc(nys + 1, 8); s->psiHyz = calloc(nys + 1, 8);
    s->psiDyx = calloc(nys + 1, 8); s->psiDyz = calloc(nys + 1, 8);
    s->psiHzx = calloc(ncell, 8); s->psiHzy = calloc(ncell, 8);
    s->psiDzx = calloc(ncell, 8); s->psiDzy = calloc(ncell, 8);
    double deltas[3] = {g->dx, g->dy, g->dz};
    for (int d = 0; d < 3; d++) {
        s->pml_s[d] = calloc(L + 1, 8); s->pml_k[d] = calloc(L + 1, 8); s->pml_a[d] = calloc(L + 1, 8);
        if (L > 0)
            orc_pml_profile(L, deltas[d], g->eps0, g->mu0, g->pml_reflection, g->pml_order,
                            g->kappa_max, g->alpha_max, s->pml_s[d], s->pml_k[d], s->pml_a[d]);
    }
    s->nmat = nmat;
    s->mats = (orc_mat *)calloc(nmat, sizeof(orc_mat));
    for (int m = 0; m < nmat; m++) {
        orc_mat *om = &s->mats[m];
        om->m = mats[m];
        int n = mats[m].natom;
        if (n > 0) {
            om->qezv = malloc(9 * n * 8); memcpy(om->qezv, mats[m].qezv, 9 * n * 8);
            om->qezm = malloc(9 * n * 8); memcpy(om->qezm, mats[m].qezm, 9 * n * 8);
            om->dmm = malloc(9 * n * n * 8); memcpy(om->dmm, mats[m].dmm, 9 * n * n * 8);
        }
        om->m.qezv = om->qezv; om->m.qezm = om->qezm; om->m.dmm = om->dmm;
        if (orc_eps_inverse(mats[m].eps_re, mats[m].eps_im, om->inv_re, om->inv_im) != 0) {
            set_err(err, errlen, "Error: Singular dielectric tensor!");
            orc_destroy(s);
            return NULL;
        }
        if (mats[m].use_ade) {
            double gama[9], gp[9], gm_inv[9];
            int rc = orc_ade_coefs(mats[m].eps_re, mats[m].eps_im, g->f_source, g->dt, gama,
                                   om->wp2, gp, om->gm, om->gp_inv, gm_inv);
            if (rc == -1) { set_err(err, errlen, "Error: EPS_INF is too small"); orc_destroy(s); return NULL; }
            if (rc == -2) { set_err(err, errlen, "Error: epsr1 can not be inversed"); orc_destroy(s); return NULL; }
        }
    }
    s->medium = (int32_t *)malloc(ncell * sizeof(int32_t));
    memcpy(s->medium, medium, ncell * sizeof(int32_t));
    s->ion_off = (int64_t *)malloc(ncell * sizeof(int64_t));
    s->ade_off = (int64_t *)malloc(ncell * sizeof(int64_t));
    int64_t nion = 0, nade = 0;
    for (size_t c = 0; c < ncell; c++) {
        int id = medium[c];
        if (id < 0 || id >= nmat) {
            set_err(err, errlen, "medium id out of range");
            orc_destroy(s);
            return NULL;
        }
        int n = mats[id].natom;
        s->ion_off[c] = n > 0 ? nion : -1;
        nion += 3 * n;
        s->ade_off[c] = mats[id].use_ade ? nade : -1;
        nade += mats[id].use_ade ? 3 : 0;
    }
    s->r = calloc(nion + 1, 8); s->v = calloc(nion + 1, 8);
    s->P = calloc(nade + 1, 8); s->Po = calloc(nade + 1, 8);
    return s;
}

void orc_set_sources(orc_sim *s, int n, const phpfdtd_source *src) {
    free(s->src);
    s->nsrc = n;
    s->src = (phpfdtd_source *)malloc((n + 1) * sizeof(phpfdtd_source));
    if (n) memcpy(s->src, src, n * sizeof(phpfdtd_source));
}

/* One reference time step t (fdtd.c:3116-3531, single rank). */
static void orc_one_step(orc_sim *s) {
    const phpfdtd_grid *g = &s->g;
    const int nx = s->nx, ny = s->ny, nz = s->nz, L = s->L;
    const double dt = g->dt, eps0 = g->eps0, mu0 = g->mu0;
    const double dx = g->dx, dy = g->dy, dz = g->dz;
    double *Ex = s->F[0], *Ey = s->F[1], *Ez = s->F[2];
    double *Dx = s->F[3], *Dy = s->F[4], *Dz = s->F[5];
    double *Hx = s->F[6], *Hy = s->F[7], *Hz = s->F[8];
    const double t_now = s->step * dt; /* fdtd.c:3117 */

#define PSIX(a, ii, j, k) a[(((size_t)(ii)) * ny + (j)) * nz + (k)]
#define PSIY(a, i, jj, k) a[(((size_t)(i)) * (2 * L) + (jj)) * nz + (k)]
#define PSIZ(a, i, j, k) a[CI(s, i, j, k)]
#define PML_BC(dir, idx, kap, bb, cc)                                               \
    do {                                                                           \
        double sg = s->pml_s[dir][idx];                                            \
        kap = s->pml_k[dir][idx];                                                  \
        double al = s->pml_a[dir][idx];                                            \
        double aw = sg / (kap * eps0) + al / eps0;                                 \
        bb = exp(-aw * dt);                                                        \
        cc = (aw > 1e-30) ? sg / (kap * kap * eps0) * (bb - 1.0) / aw : 0.0;       \
    } while (0)

    /* 1. H update, fdtd.c:3123-3231.  Each (i, j, k) writes only its own H and
     * psi and reads E, which this loop does not write, so the i iterations
     * are independent and every value is computed exactly as in a serial run. */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx - 1; i++)
        for (int j = 0; j < ny - 1; j++)
            for (int k = 0; k < nz; k++) {
                double pxy = 0.0, pxz = 0.0, pyx = 0.0, pyz = 0.0, pzx = 0.0, pzy = 0.0;
                double bx = 1.0, cx = dt, by = 1.0, cy = dt, bz = 1.0, cz = dt;
                double kx = 1.0, ky = 1.0, kz = 1.0;
                size_t c = FI(s, i, j, k);
                double dEy_dz = (Ey[FI(s, i, j, k + 1)] - Ey[c]) / dz;
                double dEz_dy = (Ez[FI(s, i, j + 1, k)] - Ez[c]) / dy;
                double dEz_dx = (Ez[FI(s, i + 1, j, k)] - Ez[c]) / dx;
                double dEx_dz = (Ex[FI(s, i, j, k + 1)] - Ex[c]) / dz;
                double dEx_dy = (Ex[FI(s, i, j + 1, k)] - Ex[c]) / dy;
                double dEy_dx = (Ey[FI(s, i + 1, j, k)] - Ey[c]) / dx;
                if (L > 0 && i < L) {
                    PML_BC(0, i, kx, bx, cx);
                    pxy = PSIX(s->psiHxy, i, j, k);
                    pxz = PSIX(s->psiHxz, i, j, k);
                    PSIX(s->psiHxy, i, j, k) = bx * pxy + cx * (dEz_dx / mu0);
                    PSIX(s->psiHxz, i, j, k) = bx * pxz + cx * (-dEy_dx / mu0);
                } else if (L > 0 && i >= nx - L) {
                    int idx = nx - 1 - i;
                    PML_BC(0, idx, kx, bx, cx);
                    pxy = PSIX(s->psiHxy, L + idx, j, k);
                    pxz = PSIX(s->psiHxz, L + idx, j, k);
                    PSIX(s->psiHxy, L + idx, j, k) = bx * pxy + cx * (dEz_dx / mu0);
                    PSIX(s->psiHxz, L + idx, j, k) = bx * pxz + cx * (-dEy_dx / mu0);
                }
                if (L > 0 && j < L) {
                    PML_BC(1, j, ky, by, cy);
                    pyx = PSIY(s->psiHyx, i, j, k);
                    pyz = PSIY(s->psiHyz, i, j, k);
                    PSIY(s->psiHyx, i, j, k) = by * pyx + cy * (-dEz_dy / mu0);
                    PSIY(s->psiHyz, i, j, k) = by * pyz + cy * (dEx_dy / mu0);
                } else if (j >= ny - L) {
                    int idx = ny - 1 - j;
                    PML_BC(1, idx, ky, by, cy);
                    pyx = PSIY(s->psiHyx, i, L + idx, k);
                    pyz = PSIY(s->psiHyz, i, L + idx, k);
                    PSIY(s->psiHyx, i, L + idx, k) = by * pyx + cy * (-dEz_dy / mu0);
                    PSIY(s->psiHyz, i, L + idx, k) = by * pyz + cy * (dEx_dy / mu0);
                }
                int inz = (k < L || k >= nz - L);
                if (inz) {
                    int idx = k < L ? k : nz - 1 - k;
                    PML_BC(2, idx, kz, bz, cz);
                    pzx = PSIZ(s->psiHzx, i, j, k);
                    pzy = PSIZ(s->psiHzy, i, j, k);
                }
                Hx[c] += dt * ((-dEz_dy / mu0) / ky + pyx + (dEy_dz / mu0) / kz + pzx);
                Hy[c] += dt * ((dEz_dx / mu0) / kx + pxy + (-dEx_dz / mu0) / kz + pzy);
                Hz[c] += dt * ((-dEy_dx / mu0) / kx + pxz + (dEx_dy / mu0) / ky + pyz);
                if (inz) {
                    PSIZ(s->psiHzx, i, j, k) = bz * pzx + cz * (dEy_dz / mu0);
                    PSIZ(s->psiHzy, i, j, k) = bz * pzy + cz * (-dEx_dz / mu0);
                }
            }

    /* 2. D update, fdtd.c:3238-3356 */
    const double wloss = dt * g->f_source * 2.0 * 3.14159265; /* fdtd.c:3350 */
    /* own D and psi written, H and E only read: independent i iterations */
#pragma omp parallel for schedule(static)
    for (int i = 1; i < nx - 1; i++)
        for (int j = 1; j < ny - 1; j++)
            for (int k = 0; k < nz; k++) {
                double pxy = 0.0, pxz = 0.0, pyx = 0.0, pyz = 0.0, pzx = 0.0, pzy = 0.0;
                double bx = 1.0, cx = dt, by = 1.0, cy = dt, bz = 1.0, cz = dt;
                double kx = 1.0, ky = 1.0, kz = 1.0;
                size_t c = FI(s, i, j, k);
                double dHz_dy = (Hz[c] - Hz[FI(s, i, j - 1, k)]) / dy;
                double dHy_dz = (Hy[c] - Hy[FI(s, i, j, k - 1)]) / dz;
                double dHx_dz = (Hx[c] - Hx[FI(s, i, j, k - 1)]) / dz;
                double dHz_dx = (Hz[c] - Hz[FI(s, i - 1, j, k)]) / dx;
                double dHy_dx = (Hy[c] - Hy[FI(s, i - 1, j, k)]) / dx;
                double dHx_dy = (Hx[c] - Hx[FI(s, i, j - 1, k)]) / dy;
                if (L > 0 && i < L) {
                    PML_BC(0, i, kx, bx, cx);
                    pxy = PSIX(s->psiDxy, i, j, k);
                    pxz = PSIX(s->psiDxz, i, j, k);
                    PSIX(s->psiDxy, i, j, k) = bx * pxy + cx * (-dHz_dx);
                    PSIX(s->psiDxz, i, j, k) = bx * pxz + cx * (dHy_dx);
                } else if (L > 0 && i >= nx - L) {
                    int idx = nx - 1 - i;
                    PML_BC(0, idx, kx, bx, cx);
                    pxy = PSIX(s->psiDxy, L + idx, j, k);
                    pxz = PSIX(s->psiDxz, L + idx, j, k);
                    PSIX(s->psiDxy, L + idx, j, k) = bx * pxy + cx * (-dHz_dx);
                    PSIX(s->psiDxz, L + idx, j, k) = bx * pxz + cx * (dHy_dx);
                }
                if (L > 0 && j < L) {
                    PML_BC(1, j, ky, by, cy);
                    pyx = PSIY(s->psiDyx, i, j, k);
                    pyz = PSIY(s->psiDyz, i, j, k);
                    PSIY(s->psiDyx, i, j, k) = by * pyx + cy * (dHz_dy);
                    PSIY(s->psiDyz, i, j, k) = by * pyz + cy * (-dHx_dy);
                } else if (j >= ny - L) {
                    int idx = ny - 1 - j;
                    PML_BC(1, idx, ky, by, cy);
                    pyx = PSIY(s->psiDyx, i, L + idx, k);
                    pyz = PSIY(s->psiDyz, i, L + idx, k);
                    PSIY(s->psiDyx, i, L + idx, k) = by * pyx + cy * (dHz_dy);
                    PSIY(s->psiDyz, i, L + idx, k) = by * pyz + cy * (-dHx_dy);
                }
                int inz = (k < L || k >= nz - L);
                if (inz) {
                    int idx = k < L ? k : nz - 1 - k;
                    PML_BC(2, idx, kz, bz, cz);
                    pzx = PSIZ(s->psiDzx, i, j, k);
                    pzy = PSIZ(s->psiDzy, i, j, k);
                }
                Dx[c] += dt * ((dHz_dy) / ky + pyx + (-dHy_dz) / kz + pzx);
                Dy[c] += dt * ((-dHz_dx) / kx + pxy + (dHx_dz) / kz + pzy);
                Dz[c] += dt * ((dHy_dx) / kx + pxz + (-dHx_dy) / ky + pyz);
                if (inz) {
                    PSIZ(s->psiDzx, i, j, k) = bz * pzx + cz * (-dHy_dz);
                    PSIZ(s->psiDzy, i, j, k) = bz * pzy + cz * (dHx_dz);
                }
                const orc_mat *om = &s->mats[s->medium[CI(s, i, j, k)]];
                if (om->m.use_ade == 0) {
                    const double *ei = om->m.eps_im;
                    Dx[c] -= wloss * (ei[0] * Ex[c] + ei[1] * Ey[c] + ei[2] * Ez[c]) * eps0;
                    Dy[c] -= wloss * (ei[3] * Ex[c] + ei[4] * Ey[c] + ei[5] * Ez[c]) * eps0;
                    Dz[c] -= wloss * (ei[6] * Ex[c] + ei[7] * Ey[c] + ei[8] * Ez[c]) * eps0;
                }
            }

    /* 3. sources, fdtd.c:3363-3396 */
    for (int q = 0; q < s->nsrc; q++) {
        const phpfdtd_source *sr = &s->src[q];
        double value = orc_source_value(sr, t_now);
        int i = sr->x, j = sr->y, k = sr->z;
        if (k < 0 || k >= nz || i < 0 || i >= nx || j < 0 || j >= ny) continue;
        size_t c = FI(s, i, j, k);
        if (sr->component == 'x') Dx[c] += value * dt;
        else if (sr->component == 'y') Dy[c] += value * dt;
        else if (sr->component == 'z') Dz[c] += value * dt;
    }

    /* 4. D -> E, ADE, ion EOM, fdtd.c:3404-3531: every cell reads and writes
     * only its own E, D, P and r, v: independent i iterations */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                size_t cc = CI(s, i, j, k), c = FI(s, i, j, k);
                const orc_mat *om = &s->mats[s->medium[cc]];
                const phpfdtd_material *m = &om->m;
                int natom = m->natom;
                double *r = natom ? s->r + s->ion_off[cc] : NULL;
                double *v = natom ? s->v + s->ion_off[cc] : NULL;
                double lf = 1.0 / (1.0 + m->sigma * dt / (2 * eps0));
                double lt = (m->sigma * dt / (2 * eps0)) * lf;
                double Pr[3];
                for (int e = 0; e < 3; e++) {
                    Pr[e] = 0.0;
                    for (int a = 0; a < natom; a++)
                        for (int d = 0; d < 3; d++) Pr[e] += m->qezv[a * 9 + e * 3 + d] * r[a * 3 + d];
                }
                double Ea, Eb, Ec, Da, Db, Dc;
                if (fabs(m->anglez) > 1e-10) {
                    double cs = cos(m->anglez), sn = sin(m->anglez);
                    Ea = Ex[c] * cs - Ey[c] * sn;
                    Eb = Ex[c] * sn + Ey[c] * cs;
                    Ec = Ez[c];
                    Da = Dx[c] * cs - Dy[c] * sn;
                    Db = Dx[c] * sn + Dy[c] * cs;
                    Dc = Dz[c];
                } else if (fabs(m->anglex) > 1e-10) {
                    double cs = cos(-m->anglex), sn = sin(-m->anglex);
                    Ea = Ex[c];
                    Eb = Ey[c] * cs - Ez[c] * sn;
                    Ec = Ey[c] * sn + Ez[c] * cs;
                    Da = Dx[c];
                    Db = Dy[c] * cs - Dz[c] * sn;
                    Dc = Dy[c] * sn + Dz[c] * cs;
                } else {
                    Ea = Ex[c]; Eb = Ey[c]; Ec = Ez[c];
                    Da = Dx[c]; Db = Dy[c]; Dc = Dz[c];
                }
                double Dav = Da - Pr[0], Dbv = Db - Pr[1], Dcv = Dc - Pr[2];
                double Ean, Ebn, Ecn;
                if (m->use_ade == 0) {
                    const double *iv = om->inv_re;
                    Ean = lf * (iv[0] * Dav + iv[1] * Dbv + iv[2] * Dcv) / eps0;
                    Ebn = lf * (iv[3] * Dav + iv[4] * Dbv + iv[5] * Dcv) / eps0;
                    Ecn = lf * (iv[6] * Dav + iv[7] * Dbv + iv[8] * Dcv) / eps0;
                } else {
                    double *P = s->P + s->ade_off[cc], *Po = s->Po + s->ade_off[cc];
                    double tv[3], Pn[3];
                    for (int id = 0; id < 3; id++) {
                        tv[id] = 2.0 * P[id];
                        tv[id] += -om->gm[id * 3 + 0] * Po[0] - om->gm[id * 3 + 1] * Po[1] -
                                  om->gm[id * 3 + 2] * Po[2] +
                                  eps0 * dt * dt *
                                      (om->wp2[id * 3 + 0] * Ea + om->wp2[id * 3 + 1] * Eb +
                                       om->wp2[id * 3 + 2] * Ec);
                    }
                    for (int id = 0; id < 3; id++) {
                        Pn[id] = 0.0;
                        for (int kk = 0; kk < 3; kk++) Pn[id] += om->gp_inv[id * 3 + kk] * tv[kk];
                    }
                    for (int id = 0; id < 3; id++) { Po[id] = P[id]; P[id] = Pn[id]; }
                    Ean = (Dav - P[0]) / eps0 / ORC_EPS_INF;
                    Ebn = (Dbv - P[1]) / eps0 / ORC_EPS_INF;
                    Ecn = (Dcv - P[2]) / eps0 / ORC_EPS_INF;
                }
                Ea = Ean - lt * Ean;
                Eb = Ebn - lt * Ebn;
                Ec = Ecn - lt * Ecn;
                if (fabs(m->anglez) > 1e-10) {
                    double cs = cos(m->anglez), sn = sin(m->anglez);
                    Ex[c] = Ea * cs + Eb * sn;
                    Ey[c] = -Ea * sn + Eb * cs;
                    Ez[c] = Ec;
                } else if (fabs(m->anglex) > 1e-10) {
                    double cs = cos(-m->anglex), sn = sin(-m->anglex);
                    Ex[c] = Ea;
                    Ey[c] = Eb * cs + Ec * sn;
                    Ez[c] = -Eb * sn + Ec * cs;
                } else {
                    Ex[c] = Ea; Ey[c] = Eb; Ez[c] = Ec;
                }
                if (natom > 0) {
                    double acc[3 * 64];
                    for (int a = 0; a < natom; a++)
                        for (int d = 0; d < 3; d++) {
                            double av = m->qezm[a * 9 + 0 * 3 + d] * Ea + m->qezm[a * 9 + 1 * 3 + d] * Eb +
                                        m->qezm[a * 9 + 2 * 3 + d] * Ec;
                            for (int ja = 0; ja < natom; ja++)
                                for (int jd = 0; jd < 3; jd++)
                                    av -= m->dmm[a * 3 * natom * 3 + d * natom * 3 + ja * 3 + jd] *
                                          r[ja * 3 + jd];
                            acc[a * 3 + d] = av;
                        }
                    for (int q = 0; q < 3 * natom; q++) {
                        r[q] += v[q] * dt + 0.5 * acc[q] * dt * dt;
                        v[q] += acc[q] * dt;
                    }
                }
            }
    s->step++;
#undef PSIX
#undef PSIY
#undef PSIZ
#undef PML_BC
}

int orc_step(orc_sim *s, int n) {
    for (int t = 0; t < n; t++) orc_one_step(s);
    return 0;
}

long orc_get_step(const orc_sim *s) { return s->step; }

/* comp 0..8 in file order [nx][ny][nz]; k may include the halo when kplus=1
 * (not used); out has nx*ny*nz entries. */
void orc_get_field(const orc_sim *s, int comp, double *out) {
    for (int i = 0; i < s->nx; i++)
        for (int j = 0; j < s->ny; j++)
            for (int k = 0; k < s->nz; k++) out[CI(s, i, j, k)] = s->F[comp][FI(s, i, j, k)];
}

void orc_set_field(orc_sim *s, int comp, const double *in) {
    for (int i = 0; i < s->nx; i++)
        for (int j = 0; j < s->ny; j++)
            for (int k = 0; k < s->nz; k++) s->F[comp][FI(s, i, j, k)] = in[CI(s, i, j, k)];
}

/* Ion / ADE state of all cells of material mat, in file order
 * (kind 0 r, 1 v, 2 P, 3 Po).  Returns the number of values; with buf NULL it
 * only counts them. */
long orc_state(orc_sim *s, int mat, int kind, double *buf, int write) {
    long n = 0;
    size_t ncell = (size_t)s->nx * s->ny * s->nz;
    int natom = s->mats[mat].m.natom;
    for (size_t c = 0; c < ncell; c++) {
        if (s->medium[c] != mat) continue;
        double *p;
        int w;
        if (kind < 2) {
            if (natom == 0) continue;
            p = (kind == 0 ? s->r : s->v) + s->ion_off[c];
            w = 3 * natom;
        } else {
            if (!s->mats[mat].m.use_ade) continue;
            p = (kind == 2 ? s->P : s->Po) + s->ade_off[c];
            w = 3;
        }
        for (int q = 0; q < w; q++) {
            if (buf) {
                if (write) p[q] = buf[n];
                else buf[n] = p[q];
            }
            n++;
        }
    }
    return n;
}

/* Monitor row after the last step, fdtd.c:3806-3851: fields at z+1 (the
 * reference indexes halo-offset pointers with k+halo), r/v at z. */
void orc_monitor_row(const orc_sim *s, int x, int y, int z, int atom, int dir, double row[12]) {
    row[0] = (s->step - 1) * s->g.dt;
    for (int q = 0; q < 9; q++) row[1 + q] = s->F[q][FI(s, x, y, z + 1)];
    row[10] = row[11] = 0.0;
    size_t cc = CI(s, x, y, z);
    int natom = s->mats[s->medium[cc]].m.natom;
    if (atom < natom) {
        row[10] = s->r[s->ion_off[cc] + atom * 3 + dir];
        row[11] = s->v[s->ion_off[cc] + atom * 3 + dir];
    }
}

/* max|E|, max|Ex|, max|r| and its first argmax in (i,j,k,atom,dir) order
 * (fdtd.c:3537-3592).  idx receives i,j,k,atom,dir. */
void orc_diag(const orc_sim *s, double out[3], int idx[5]) {
    double me = 0.0, mex = 0.0, mr = 0.0;
    idx[0] = idx[1] = idx[2] = idx[3] = idx[4] = -1;
    for (int i = 0; i < s->nx; i++)
        for (int j = 0; j < s->ny; j++)
            for (int k = 0; k < s->nz; k++) {
                size_t c = FI(s, i, j, k);
                double ex = s->F[0][c], ey = s->F[1][c], ez = s->F[2][c];
                double e = sqrt(ex * ex + ey * ey + ez * ez);
                if (e > me) me = e;
                if (fabs(ex) > mex) mex = fabs(ex);
                size_t cc = CI(s, i, j, k);
                int natom = s->mats[s->medium[cc]].m.natom;
                for (int a = 0; a < natom; a++)
                    for (int d = 0; d < 3; d++) {
                        double rv = fabs(s->r[s->ion_off[cc] + a * 3 + d]);
                        if (rv > mr) {
                            mr = rv;
                            idx[0] = i; idx[1] = j; idx[2] = k; idx[3] = a; idx[4] = d;
                        }
                    }
            }
    out[0] = me; out[1] = mex; out[2] = mr;
}

/* float32 slice in the reference layout (fdtd.c:1051-1059, 1121-1126,
 * 1314-1319); the z slice holds plane pos+1.  comp 9 = |E|. */
int orc_slice(const orc_sim *s, int comp, char axis, int pos, float *out) {
    int nx = s->nx, ny = s->ny, nz = s->nz;
#define VAL(i, j, k)                                                                              \
    (comp == 9 ? sqrt(s->F[0][FI(s, i, j, k)] * s->F[0][FI(s, i, j, k)] +                         \
                      s->F[1][FI(s, i, j, k)] * s->F[1][FI(s, i, j, k)] +                         \
                      s->F[2][FI(s, i, j, k)] * s->F[2][FI(s, i, j, k)])                          \
               : s->F[comp][FI(s, i, j, k)])
    if (axis == 'z') {
        if (pos < 0 || pos >= nz) return -1;
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < ny; j++) {
                /* |E| is only defined inside the domain; the halo plane holds 0 */
                double v = (pos + 1 >= nz && comp == 9) ? 0.0 : VAL(i, j, pos + 1);
                out[i * ny + j] = (float)v;
            }
    } else if (axis == 'x') {
        if (pos < 0 || pos >= nx) return -1;
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) out[j * nz + k] = (float)VAL(pos, j, k);
    } else if (axis == 'y') {
        if (pos < 0 || pos >= ny) return -1;
        for (int i = 0; i < nx; i++)
            for (int k = 0; k < nz; k++) out[i * nz + k] = (float)VAL(i, pos, k);
    } else
        return -1;
#undef VAL
    return 0;
}

/* r and v of every atom of the cell (x,y,z), fdtd.c:3766-3775.  Returns natom. */
int orc_cell_ions(const orc_sim *s, int x, int y, int z, double *r, double *v) {
    size_t cc = CI(s, x, y, z);
    int natom = s->mats[s->medium[cc]].m.natom;
    for (int q = 0; q < 3 * natom; q++) {
        r[q] = s->r[s->ion_off[cc] + q];
        v[q] = s->v[s->ion_off[cc] + q];
    }
    return natom;
}
