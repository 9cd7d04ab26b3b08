/*
 * phpfdtd.h - C ABI of libphpfdtd_hip, the MI355X (gfx950) per-timestep
 * update of the PhP_FDTD solver.
 *
 * The reference has no operator or plugin layer: one time step is inline code
 * in main() (fdtd.c:3116-3531) working on double*** arrays.  This header
 * defines the boundary at exactly the loop bodies that code replaces:
 *
 *   fdtd.c:3120-3231  E halo + H update with CPML psi      -> phpfdtd_run
 *   fdtd.c:3235-3356  H halo + D update with CPML psi/loss -> phpfdtd_run
 *   fdtd.c:3358-3400  soft source injection                -> phpfdtd_set_sources + phpfdtd_run
 *   fdtd.c:3403-3531  D->E, tensor ADE, ion EOM            -> phpfdtd_run
 *   fdtd.c:3535-3668  max|E|, max|Ex|, max|r| + guard      -> phpfdtd_sync
 *   fdtd.c:3670-3726, 999-1400  slice sampling             -> phpfdtd_read_slice
 *   fdtd.c:3728-3883  monitor / full-monitor sampling      -> phpfdtd_set_monitor,
 *                                                             phpfdtd_read_monitor,
 *                                                             phpfdtd_read_point,
 *                                                             phpfdtd_read_ions
 *   fdtd.c:555-625, 2478-2528, 962-989  host precompute    -> phpfdtd_create
 *   frequency_analysis.c:362-390, 548-569  slice FFT bin    -> phpfdtd_set_dft,
 *                                                             phpfdtd_read_dft
 *
 * Conventions
 *   - Every function returns 0 or a negative PHPFDTD_E* code, and never calls
 *     exit()/abort().  phpfdtd_last_error() returns the message of the last
 *     failure on the calling thread.
 *   - Input pointers are borrowed for the duration of the call and deep-copied.
 *     Output buffers are allocated by the caller.  The context owns all device
 *     memory and one HIP stream.
 *   - A context is not thread-safe.  Distinct contexts may be driven from
 *     different host threads.
 *   - Arrays exchanged with the host use the reference's file order: x slowest,
 *     z fastest, i.e. element (i,j,k) sits at (i*ny + j)*nz + k, the order of
 *     medium_map.bin (fdtd.c:809-812) and of the reference's field arrays
 *     (fdtd.c:261-284).
 */
#ifndef PHPFDTD_H
#define PHPFDTD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PHPFDTD_ABI_VERSION 1

/* Return codes. */
#define PHPFDTD_OK          0
#define PHPFDTD_EINVAL     -1  /* bad argument or configuration                 */
#define PHPFDTD_ENOMEM_HBM -2  /* device allocation failed                      */
#define PHPFDTD_EHIP       -3  /* HIP runtime error                             */
#define PHPFDTD_EDIVERGED  -4  /* max|r| > 1e-10 m at a sync (fdtd.c:3626)      */
#define PHPFDTD_EFORMAT    -5  /* malformed input data                          */
#define PHPFDTD_ESTATE     -6  /* call not valid in the context's current state */
#define PHPFDTD_ENOMEM_HOST -7 /* host allocation failed                        */

/* Numerical forms. */
#define PHPFDTD_FP64 0  /* plain form: everything in FP64                          */
#define PHPFDTD_FP32 1  /* precision form: fields, CPML, ADE in FP32; DMM.r on FP32
                           MFMA with FP32 accumulation; ion r,v integrated in FP64
                           and stored as compensated FP32 (an FP32 word plus a
                           4-bit compensation code: 28 significant bits)         */
#define PHPFDTD_FP32_RV64 2  /* precision form with ion r,v stored in FP64      */
/* Both precision forms update cells of plain (non-ADE, non-lattice)
 * materials in E-form: E^{n+1} = E^n + M.dD, the same update as the
 * reference's D^{n+1} = D^n + dD, E = M.D^{n+1} (fdtd.c:3238-3462), without
 * storing D.  The D those cells report (read_field, read_point, read_slice,
 * monitor rows, the DFT) is M^-1 E; a D written to them is not used (their
 * state is E).  ADE and lattice cells, and materials whose M has no
 * inverse, keep D as the reference does. */

/* Field components (phpfdtd_read_field / write_field / read_slice). */
#define PHPFDTD_EX 0
#define PHPFDTD_EY 1
#define PHPFDTD_EZ 2
#define PHPFDTD_DX 3
#define PHPFDTD_DY 4
#define PHPFDTD_DZ 5
#define PHPFDTD_HX 6
#define PHPFDTD_HY 7
#define PHPFDTD_HZ 8
#define PHPFDTD_EMAG 9  /* |E|, slices only (fdtd.c:3693-3703) */

/* Grid, time step and CPML description.  Replaces the Config fields that the
 * hot loop reads (fdtd.c:136-169) and the init_pml() constants (fdtd.c:975-977). */
typedef struct phpfdtd_grid {
    int32_t nx, ny, nz;       /* cells                                              */
    int32_t tsteps;           /* steps covered by the source table                  */
    double  dx, dy, dz;       /* m (delta_x/y/z)                                    */
    double  dt;               /* s, already CFL-clamped (fdtd.c:1906-1910)          */
    double  eps0, mu0;        /* 8.854e-12, 4*pi*1e-7 in the reference (1751-1752)  */
    int32_t pml_thickness;    /* L                                                  */
    int32_t precision;        /* PHPFDTD_FP64, PHPFDTD_FP32 or PHPFDTD_FP32_RV64    */
    double  pml_reflection;   /* R                                                  */
    double  pml_order;        /* m, 3 in the reference                              */
    double  kappa_max;        /* 7 in the reference                                 */
    double  alpha_max;        /* 0.05 in the reference                              */
    double  f_source;         /* Hz, unique_source_frequency (fdtd.c:1774)          */
    int32_t flags;            /* reserved, 0                                        */
    int32_t reserved;
} phpfdtd_grid;

/* One material after the reference's host loading: eps from keys, table or
 * OUTCAR, and switchxz already applied to eps/QeZV/QeZM/DMM
 * (fdtd.c:2248-2470).  Mirrors the fields of Material (fdtd.c:82-113) that the
 * time step reads; the library derives eps_inv (fdtd.c:555-625), the ADE
 * coefficients (fdtd.c:2478-2528) and the rotation itself, in FP64. */
typedef struct phpfdtd_material {
    double  eps_re[9];        /* row-major xx xy xz yx yy yz zx zy zz (crystal frame) */
    double  eps_im[9];
    double  sigma;            /* S/m                                               */
    double  anglez, anglex;   /* rad; anglez wins when both are set (fdtd.c:3429)  */
    int32_t use_ade;          /* 0 or 1                                            */
    int32_t natom;            /* 0 = no lattice                                    */
    const double *qezv;       /* [natom][3 (E comp)][3 (dir)]  C/m^3  (9*natom)     */
    const double *qezm;       /* [natom][3 (E comp)][3 (dir)]  C/kg   (9*natom)     */
    const double *dmm;        /* [3*natom][3*natom] row-major, s^-2 (9*natom^2)    */
} phpfdtd_material;

/* One soft source (Source, fdtd.c:123-133).  Values follow
 * calculate_source_value (fdtd.c:1923-1947) and are evaluated on the host in
 * FP64 for every step t at time t*dt. */
typedef struct phpfdtd_source {
    int32_t type;             /* 0 gaussian, 1 sinusoidal, 2 ricker, 3 dipole      */
    int32_t x, y, z;
    int32_t component;        /* 'x', 'y' or 'z'                                   */
    int32_t reserved;
    double  amplitude, frequency, t0, tau, start_time, end_time;
} phpfdtd_source;

/* Diagnostics of fdtd.c:3535-3668. */
typedef struct phpfdtd_diag {
    int64_t step;             /* steps completed so far                            */
    double  max_e;            /* max |E| over all cells                            */
    double  max_ex;           /* max |Ex|                                          */
    double  max_r;            /* max |r| over all ion dofs (m)                     */
    int32_t r_i, r_j, r_k, r_atom, r_dir;  /* first argmax in (i,j,k,atom,dir) order; -1 if none */
    int32_t diverged;         /* max_r > 1e-10                                     */
} phpfdtd_diag;

/* Per-kernel timing from phpfdtd_profile. */
typedef struct phpfdtd_ktime {
    char    name[48];
    double  avg_ms;           /* mean duration of one launch                       */
    double  bytes;            /* algorithmic HBM bytes of one launch               */
    double  flops;            /* algorithmic flops of one launch                   */
    int32_t launches;         /* launches per step                                 */
    int32_t reserved;
} phpfdtd_ktime;

typedef struct phpfdtd_ctx phpfdtd_ctx;

/* Build a context: validates the inputs, runs the FP64 host precompute,
 * allocates the device state (all zero, as fdtd.c:3053-3100) and uploads it.
 * medium: nx*ny*nz int32 material ids in file order (fdtd.c:718-959).
 * Replaces fdtd.c:2433-2530 (coefficients), 2881-3100 (state). */
int phpfdtd_create(const phpfdtd_grid *grid, const phpfdtd_material *mats, int nmat,
                   const int32_t *medium, phpfdtd_ctx **out);

/* Install sources; replaces fdtd.c:3363-3396.  Values for steps [0, tsteps)
 * are evaluated on the host once. */
int phpfdtd_set_sources(phpfdtd_ctx *ctx, int n, const phpfdtd_source *src);

/* calculate_source_value (fdtd.c:1923-1947) at time t in FP64, the value the
 * library injects (times dt); for drivers writing source.txt (fdtd.c:3359-3367). */
double phpfdtd_source_value(const phpfdtd_source *src, double t);

/* Enable the per-step monitor ring (fdtd.c:3806-3883).  ref_compat=1 samples
 * E, D and H at plane z+1 as the reference does (its field pointers are
 * pre-offset by the halo, fdtd.c:274 with 3814) and r, v at plane z.  Each row
 * holds 12 doubles: t*dt, Ex, Ey, Ez, Dx, Dy, Dz, Hx, Hy, Hz, R, V; R and V
 * are 0 when the cell has no atom `atom`. */
int phpfdtd_set_monitor(phpfdtd_ctx *ctx, int x, int y, int z, int atom, int dir,
                        int ref_compat);

/* Enqueue nsteps time steps (replay of the captured step graph); returns
 * after enqueue. */
int phpfdtd_run(phpfdtd_ctx *ctx, int nsteps);

/* Wait for queued work, then compute the diagnostics of fdtd.c:3535-3596.
 * Returns PHPFDTD_EDIVERGED when max|r| > 1e-10 m (fdtd.c:3626). d may be NULL. */
int phpfdtd_sync(phpfdtd_ctx *ctx, phpfdtd_diag *d);

/* Copy one plane of a component as float32 in the reference's slice layout
 * (fdtd.c:999-1400): axis 'z' -> [nx][ny], 'x' -> [ny][nz], 'y' -> [nx][nz].
 * ref_compat=1 reproduces the z slice reading plane pos+1 (fdtd.c:1029). */
int phpfdtd_read_slice(phpfdtd_ctx *ctx, int comp, char axis, int pos, int ref_compat,
                       float *out);

/* Drain monitor rows recorded since the last call (at most max_rows). */
int phpfdtd_read_monitor(phpfdtd_ctx *ctx, double *rows, int max_rows, int *n_rows);

/* E, D, H at one cell in FP64: Ex Ey Ez Dx Dy Dz Hx Hy Hz (full-monitor use,
 * fdtd.c:3756-3765). z may equal nz (the zero upper halo plane). */
int phpfdtd_read_point(phpfdtd_ctx *ctx, int x, int y, int z, double *fields9);

/* r and v (3*natom each, m and m/s, [atom][dir]) of the ion cell (x,y,z).
 * Returns the number of atoms at that cell (0 if none). */
int phpfdtd_read_ions(phpfdtd_ctx *ctx, int x, int y, int z, double *r, double *v);

/* Near-field single-bin DFT of one plane, accumulated on the device in FP64
 * while the simulation runs: the quantity frequency_analysis.c computes from
 * the slice files (frequency_analysis.c:362-390 bin choice, 548-569 FFT),
 * without storing the slice time series.
 *   Samples: x_n = the plane after the update of step t0 + n*interval,
 *            n = 0 .. nsamples-1, where t0 = steps completed at this call
 *            (0 before the first phpfdtd_run, as the reference writes slices
 *            at t % output_interval == 0, fdtd.c:3670).
 *   Bin:     k = first argmin over i < nsamples/2 of |i*fs/N - f_hz| with
 *            fs = 1/(interval*dt), N = nsamples.
 *   Result:  X = sum_n x_n exp(-2 pi i k n / N) / N  (FFTW forward sign).
 * comp, axis, pos and ref_compat as in phpfdtd_read_slice; ref_compat=1 also
 * samples the float32 value a slice file holds.  Replaces a previous DFT and
 * zeroes its accumulator.  Adds 16 bytes per plane cell and per sample to the
 * context's device memory.  *bin (may be NULL) receives k. */
int phpfdtd_set_dft(phpfdtd_ctx *ctx, int comp, char axis, int pos, int ref_compat, int interval,
                    int nsamples, double f_hz, int *bin);

/* phpfdtd_set_dft without device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, bin and results as phpfdtd_set_dft, but the
 * weights and the accumulator live in a static per-device workspace of the
 * library (17 MiB on each device it runs on, held until the process ends)
 * instead of context allocations; nothing is allocated or freed.  Limits:
 * nsamples <= 65536 and a plane of at most 1048576 cells (a 1024 x 1024 z
 * plane); larger requests return PHPFDTD_EINVAL.  Every call on a device
 * shares that device's workspace and clears it: a DFT installed by this call
 * is valid until the next call of it on the same device, so uses that
 * overlap on one device (two contexts, or two threads) race. */
int phpfdtd_set_dft_noalloc(phpfdtd_ctx *ctx, int comp, char axis, int pos, int ref_compat, int interval,
                            int nsamples, double f_hz, int *bin);

/* Real and imaginary parts of X in the slice layout of phpfdtd_read_slice
 * (re or im may be NULL); *n_samples (may be NULL) receives the samples
 * accumulated so far. */
int phpfdtd_read_dft(phpfdtd_ctx *ctx, double *re, double *im, int *n_samples);

/* Batched stepping of several contexts (continuous batching for serving).
 * No counterpart in the reference, which runs one simulation per process.
 * A batch advances its member contexts together on the batch's own stream:
 * each kernel of a step (H, D/E, ADE, lattice per atom count, step epilogue)
 * is ONE launch over the tiles of all members, each tile reading its member's
 * argument block, so many small domains fill the GPU like one large one.
 * Every member computes exactly what it computes alone (same kernels and
 * arithmetic).  Members keep their own state, sources, monitor and DFT, and
 * stay readable with the single-context calls (which wait for the batch's
 * work); phpfdtd_run / phpfdtd_profile / phpfdtd_time_steps on a member
 * return PHPFDTD_ESTATE.  Membership may change between phpfdtd_batch_run
 * calls; destroying a member removes it.  Members must share the batch's
 * precision value.  A member's Yee launches take one y row per wavefront
 * from phpfdtd_batch_add on (same results).  Not thread-safe together with
 * calls on its members. */
typedef struct phpfdtd_batch phpfdtd_batch;
int  phpfdtd_batch_create(int precision, phpfdtd_batch **out);
int  phpfdtd_batch_add(phpfdtd_batch *b, phpfdtd_ctx *ctx);
int  phpfdtd_batch_remove(phpfdtd_batch *b, phpfdtd_ctx *ctx);
/* Enqueue nsteps steps of every member (each member must have nsteps left
 * before tsteps); returns after enqueue. */
int  phpfdtd_batch_run(phpfdtd_batch *b, int nsteps);
/* Wait for the batch's queued work. */
int  phpfdtd_batch_sync(phpfdtd_batch *b);
/* Wait until at most `keep` phpfdtd_batch_run calls are still queued (keep=1:
 * the latest runs on while the host prepares the next). */
int  phpfdtd_batch_wait(phpfdtd_batch *b, int keep);
int  phpfdtd_batch_size(const phpfdtd_batch *b);
void phpfdtd_batch_destroy(phpfdtd_batch *b);

/* Whole component in file order, FP64 (tests, checkpoints).  In the
 * precision forms D of plain cells reads back as M^-1 E and is not taken on
 * write (see PHPFDTD_FP32_RV64 above). */
int phpfdtd_read_field(phpfdtd_ctx *ctx, int comp, double *out);
int phpfdtd_write_field(phpfdtd_ctx *ctx, int comp, const double *in);

/* Whole state of the ion lattice / ADE polarisation in file order:
 * kind 0 = r, 1 = v (3*natom doubles per cell of material mat),
 *      2 = P_ADE, 3 = P_ADE old (3 doubles per cell of material mat).
 * In PHPFDTD_FP32, r and v are stored compensated: written values are
 * rounded to 28 significant bits (relative error <= 2^-27), and reading
 * returns the stored values exactly.
 * Cells of other materials are skipped: n_cells*width values are exchanged,
 * where n_cells is returned by phpfdtd_count_cells. */
int phpfdtd_count_cells(phpfdtd_ctx *ctx, int mat, int64_t *n_cells);
int phpfdtd_read_state(phpfdtd_ctx *ctx, int mat, int kind, double *out);
int phpfdtd_write_state(phpfdtd_ctx *ctx, int mat, int kind, const double *in);

/* Device bytes a context for these inputs would hold (admission control). */
int phpfdtd_footprint_bytes(const phpfdtd_grid *grid, const phpfdtd_material *mats, int nmat,
                            const int32_t *medium, size_t *bytes);

/* Run nsteps with HIP events around every kernel launch on the context's
 * stream, and report the mean duration of each kernel with its algorithmic
 * bytes and flops.  n_out receives the number of entries. */
int phpfdtd_profile(phpfdtd_ctx *ctx, int nsteps, phpfdtd_ktime *out, int max_out, int *n_out);

/* Mean wall time per step (ms) of nsteps graph replays, timed with HIP events
 * on the context's stream after a barrier. */
int phpfdtd_time_steps(phpfdtd_ctx *ctx, int nsteps, double *ms_per_step);

/* Steps completed so far. */
int64_t phpfdtd_step(const phpfdtd_ctx *ctx);

/* Message of the last failure on this thread ("" if none). */
const char *phpfdtd_last_error(void);

/* sizeof of phpfdtd_grid, _material, _source, _diag, _ktime as compiled into
 * the library (binding checks); returns the number of sizes known (5). */
int  phpfdtd_struct_sizes(int *sizes, int n);

int  phpfdtd_abi_version(void);
void phpfdtd_destroy(phpfdtd_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif /* PHPFDTD_H */
