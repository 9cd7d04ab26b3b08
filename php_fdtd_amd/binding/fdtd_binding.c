/*
 * fdtd_binding.c - the reference-side binding of libphpfdtd_hip: what fdtd.c's
 * main() calls when its single-rank time loop (fdtd.c:3116-3883) is handed to
 * the library.  It is written against the reference's own types (Config,
 * Material, Source: fdtd.c:72-169) and is compiled inside fdtd.c's
 * translation unit, after fdtd.c (the types live in that file, not in a
 * header):
 *
 *     #define main fdtd_reference_main   (or edit main() itself)
 *     #include "fdtd.c"
 *     #undef main
 *     #include "fdtd_binding.c"
 *
 * oracle/bound_main.c does exactly that and drives these two functions from
 * the reference's own read_config() and read_medium_map_parallel(); the GPU
 * tests run it on the reference's golden cases.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "phpfdtd.h"

/* A library context from the reference's loaded state: the Config after
 * read_config() and main()'s host material loading (fdtd.c:1730-1922,
 * 2248-2470: eps from keys, table or OUTCAR, switchxz applied), and the
 * medium as read_medium_map_parallel() returns it on one rank
 * ([nx][ny][nz + 2 halo], cell k at index k + halo, fdtd.c:718-959).
 * The CPML constants are those of init_pml (fdtd.c:975-977). */
static int phpfdtd_from_reference(const Config *cfg, int ***medium, int halo, int precision, phpfdtd_ctx **ctx) {
    phpfdtd_grid g;
    memset(&g, 0, sizeof g);
    g.nx = cfg->nx; g.ny = cfg->ny; g.nz = cfg->nz; g.tsteps = cfg->tsteps;
    g.dx = cfg->delta_x; g.dy = cfg->delta_y; g.dz = cfg->delta_z; g.dt = cfg->dt;
    g.eps0 = cfg->epsilon0; g.mu0 = cfg->mu0;
    g.pml_thickness = cfg->pml_thickness; g.pml_reflection = cfg->pml_reflection;
    g.pml_order = 3.0; g.kappa_max = 7.0; g.alpha_max = 0.05;
    g.f_source = cfg->unique_source_frequency;
    g.precision = precision;

    const int nmat = cfg->num_materials;
    phpfdtd_material *pm = (phpfdtd_material *)calloc((size_t)nmat, sizeof *pm);
    int32_t *ids = (int32_t *)malloc(sizeof(int32_t) * (size_t)cfg->nx * cfg->ny * cfg->nz);
    if (!pm || !ids) {
        free(pm);
        free(ids);
        return PHPFDTD_ENOMEM_HOST;
    }
    for (int m = 0; m < nmat; m++) {
        const Material *mt = &cfg->materials[m];
        const DielectricTensor *e = &mt->eps;
        const double re[9] = {e->xx, e->xy, e->xz, e->yx, e->yy, e->yz, e->zx, e->zy, e->zz};
        const double im[9] = {e->xxi, e->xyi, e->xzi, e->yxi, e->yyi, e->yzi, e->zxi, e->zyi, e->zzi};
        memcpy(pm[m].eps_re, re, sizeof re);
        memcpy(pm[m].eps_im, im, sizeof im);
        pm[m].sigma = mt->sigma;
        pm[m].anglez = mt->anglez;
        pm[m].anglex = mt->anglex;
        pm[m].use_ade = mt->use_ADE;
        pm[m].natom = mt->natom;
        pm[m].qezv = mt->QeZV;
        pm[m].qezm = mt->QeZM;
        pm[m].dmm = mt->DMM;
    }
    /* file order (i*ny + j)*nz + k, without the halo planes */
    for (int i = 0; i < cfg->nx; i++)
        for (int j = 0; j < cfg->ny; j++)
            for (int k = 0; k < cfg->nz; k++)
                ids[((size_t)i * cfg->ny + j) * cfg->nz + k] = medium[i][j][k + halo];

    int rc = phpfdtd_create(&g, pm, nmat, ids, ctx);
    free(ids);
    free(pm);
    if (rc) return rc;

    phpfdtd_source *s = (phpfdtd_source *)calloc((size_t)(cfg->num_sources > 0 ? cfg->num_sources : 1), sizeof *s);
    if (!s) return PHPFDTD_ENOMEM_HOST;
    for (int q = 0; q < cfg->num_sources; q++) {
        const Source *r = &cfg->sources[q];
        s[q].type = r->type;
        s[q].x = r->position[0]; s[q].y = r->position[1]; s[q].z = r->position[2];
        s[q].component = r->component;
        s[q].amplitude = r->amplitude; s[q].frequency = r->frequency;
        s[q].t0 = r->t0; s[q].tau = r->tau;
        s[q].start_time = r->start_time; s[q].end_time = r->end_time;
    }
    rc = phpfdtd_set_sources(*ctx, cfg->num_sources, s);
    free(s);
    if (rc) return rc;
    if (cfg->monitor_enabled) /* fields at plane z+1, as fdtd.c:3814 samples them */
        rc = phpfdtd_set_monitor(*ctx, cfg->monitor_x, cfg->monitor_y, cfg->monitor_z, cfg->monitor_atom,
                                 cfg->monitor_dir, 1);
    return rc;
}

/* The time loop of fdtd.c:3116-3883 on the device.  Diagnostics come at the
 * reference's cadence: after step t for every t % output_interval == 0
 * (fdtd.c:3535-3594), with its divergence report and stop (fdtd.c:3626-3667).
 * Monitor rows (t, E, D, H[, R, V]) go to `monitor` at "%.17e" (the reference
 * writes "%e"), one per step.  Returns 0, PHPFDTD_EDIVERGED or an error. */
static int phpfdtd_reference_loop(const Config *cfg, phpfdtd_ctx *ctx, FILE *monitor, int monitor_rv) {
    const int every = cfg->output_interval > 0 ? cfg->output_interval : cfg->tsteps;
    const int cap = every + 1;
    double *rows = (double *)malloc(sizeof(double) * 12 * (size_t)cap);
    if (!rows) return PHPFDTD_ENOMEM_HOST;
    int rc = 0;
    for (int t = 0; t < cfg->tsteps && !rc;) {
        /* run through the next diagnostic step t' (t' % every == 0), inclusive */
        const int next = t % every == 0 ? t : t + (every - t % every);
        int n = (next < cfg->tsteps ? next : cfg->tsteps - 1) - t + 1;
        rc = phpfdtd_run(ctx, n);
        if (rc) break;
        t += n;
        if ((t - 1) % every == 0) {
            phpfdtd_diag d;
            rc = phpfdtd_sync(ctx, &d);
            if (rc == PHPFDTD_EDIVERGED) {
                printf("ERROR: r field too large at step %d\n", t - 1);
                printf("Max r value: %e\n", d.max_r);
                printf("Location: i=%d, j=%d, k=%d\n", d.r_i, d.r_j, d.r_k);
                printf("Atom: %d, Direction: %d\n", d.r_atom, d.r_dir);
            } else if (!rc) {
                printf("it %d/%d (%.1f%%), t = %.5e s,|E|_max = %.2e V/m,r_max = %.2e A,\n", t - 1, cfg->tsteps,
                       100.0 * (t - 1) / cfg->tsteps, (t - 1) * cfg->dt, d.max_e, d.max_r / 1e-10);
                fflush(stdout);
            }
        }
        if (monitor && rc != PHPFDTD_EDIVERGED) {
            int got = 0;
            do {
                int e = phpfdtd_read_monitor(ctx, rows, cap, &got);
                if (e) { rc = e; break; }
                for (int r = 0; r < got; r++)
                    for (int q = 0; q < (monitor_rv ? 12 : 10); q++)
                        fprintf(monitor, q + 1 < (monitor_rv ? 12 : 10) ? "%.17e\t" : "%.17e\n", rows[12 * r + q]);
            } while (got > 0);
        }
    }
    free(rows);
    return rc;
}
