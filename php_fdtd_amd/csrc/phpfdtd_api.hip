// phpfdtd_api.hip - C ABI of libphpfdtd_hip (include/phpfdtd.h).
//
// A context owns the whole simulation state in HBM and one HIP stream.  One
// time step is a fixed plan of kernel launches:
//     h_update -> de_update -> [ade_update per ADE material]
//              -> [ion_update per lattice material] -> step_end
// captured once into hipGraphs of 1 and 16 steps and replayed by
// phpfdtd_run.  The step index lives on the device (advanced by step_end), so
// replays need no host work.  Sampling, diagnostics and state access run on
// the same stream outside the graphs.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <functional>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "host_precompute.h"
#include "phpfdtd_internal.h"

namespace phpfdtd {
template <typename real> void launch_yee(const YeeArgs<real> &, hipStream_t);
template <typename real> void launch_h(const YeeArgs<real> &, hipStream_t);
template <typename real> void launch_de(const YeeArgs<real> &, hipStream_t);
template <typename real> void launch_yee_batch(const BatchLaunch<YeeArgs<real>> &, int, hipStream_t);
template <typename real> void launch_h_batch(const BatchLaunch<YeeArgs<real>> &, hipStream_t);
template <typename real> void launch_de_batch(const BatchLaunch<YeeArgs<real>> &, int, hipStream_t);
template <typename real> void launch_slot_batch(const BatchLaunch<SlotArgs<real>> &, int, hipStream_t);
template <typename real> int64_t slot_blocks(const SlotArgs<real> &);
template <typename real> int slot_kind(const SlotArgs<real> &);
template <typename real> void launch_step_end_batch(const StepEndArgs *, int, hipStream_t);
template <typename real> void launch_ade(const SlotArgs<real> &, hipStream_t);
template <typename real> void launch_ion(const SlotArgs<real> &, hipStream_t);
template <typename real> void launch_step_end(const StepEndArgs &, hipStream_t);
template <typename real> void launch_dft(const DftArgs &, hipStream_t);
template <typename real> void launch_derive(const DeriveArgs &, hipStream_t);
template <typename T> void launch_fill(T *, T, int64_t, hipStream_t);
template <typename real>
void launch_reduce_e(const real *, const real *, const real *, int, int, int, int, unsigned long long *,
                     hipStream_t);
void launch_diag_r(const void *, bool, const uint32_t *, int64_t, IonLayout, int, int, int, unsigned long long *,
                   unsigned long long *,
                     hipStream_t);
template <typename real>
void launch_slice(const real *, const real *, const real *, const real *, int, char, int, int, int, int, int,
                  float *, hipStream_t);
}  // namespace phpfdtd

using namespace phpfdtd;

// Message of the last failure on this thread.  A fixed buffer, so reporting
// an error (a host allocation failure included) never allocates or throws.
static thread_local char g_err[1024];

static int fail(int code, const char *fmt, ...) noexcept {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

#define HIPCHK(x)                                                                      \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess) return fail(PHPFDTD_EHIP, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)

namespace {

struct Special {
    int id = 0;
    MatPre pre{};
    std::vector<uint32_t> cells;  // x-fastest field indices, ascending
    IonLayout lay{};              // lattice state layout r[tile][dof / w][64][w]
    std::vector<double> qezv, qezm, dmm;
    uint32_t *d_cells = nullptr;
    void *d_P = nullptr, *d_Po = nullptr;
    void *d_r = nullptr, *d_v = nullptr;        // lattice state, layout `lay`
    double *d_dmm = nullptr, *d_qezv = nullptr, *d_qezm = nullptr;
    float *d_aop = nullptr, *d_qop = nullptr;  // MFMA operand tables (precision form)
    SlotArgs<double> sa64{};                    // argument block of its slot kernel (current plan)
    SlotArgs<float> sa32{};
};

// Stages of one time step, in issue order: the Yee kernel(s), the ADE slot
// kernels, samples of D and H, the lattice kernels, the epilogue.
enum Stage { K_YEE, K_DE, K_PRE, K_ION, K_POST };

struct Kernel {
    std::string name;
    double bytes = 0, flops = 0;
    int stage = K_DE;
    std::function<void(hipStream_t)> fn;
};

// Device bytes of every allocation a context makes; shared by create and
// phpfdtd_footprint_bytes.
struct Layout {
    int nxp = 0;
    size_t ncellp = 0;
    size_t field = 0, psi_x = 0, psi_y = 0, psi_z = 0, medium = 0, small = 0, slots = 0, ring = 0;
    // E, H in two sets and D (15 field arrays); H psi in two sets and D psi
    size_t total() const { return 15 * field + 6 * (psi_x + psi_y + psi_z) + medium + small + slots + ring; }
};

}  // namespace

struct phpfdtd_ctx {
    phpfdtd_grid g{};
    int prec = 0;     // PHPFDTD_FP64 or PHPFDTD_FP32 (the field form)
    bool pk = false;  // lattice r, v in compensated FP32 (PHPFDTD_FP32), else FP64
    size_t rs = 8;
    int nx = 0, ny = 0, nz = 0, nxp = 0, L = 0;
    size_t plane = 0, ncellp = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<void *> allocs;
    size_t bytes = 0;
    // field arrays of E/H sets 0 and 1 (Ex..Hz; the D entries of set 1 are
    // those of set 0), and CPML psi slabs likewise (Hxy Hxz Dxy Dxz | Hyx Hyz
    // Dyx Dyz | Hzx Hzy Dzx Dzy; the D slabs single).  After s steps set s & 1
    // is current.
    void *F[2][9] = {};
    void *psi[2][12] = {};
    uint8_t *d_medium = nullptr;
    std::vector<uint8_t> medium;  // host copy, device layout
    int nmat = 0;
    std::vector<MatPre> pre;
    void *d_mats = nullptr;
    void *d_axis[3][5] = {};
    int *d_slot[3] = {};
    void *d_axrow[3] = {};  // y and z coefficients packed per row / plane (AxRow)
    std::vector<Special> specials;
    int special_of[kMaxMat];
    int64_t lossy_interior = 0;  // interior cells of materials with eps_im loss
    // E-form (FP32): plain cells carry E alone; materials with dstate keep D
    // (slot materials, and plain ones whose M has no inverse).  dm: M^-1 of
    // each E-form material, for reading D back.
    bool eform = false;
    std::vector<uint8_t> dstate, derive;
    std::vector<double> dm;
    uint64_t dmask = 0;
    int64_t dstate_cells = 0;
    double *d_dm = nullptr;
    uint8_t *d_derive = nullptr;
    bool fused = true;     // one-pass Yee kernel (yee_update), else h_update + de_update
    int yee_kchunk = 16;   // z planes per Yee tile
    int yee_rows = 8;      // y rows per work-group of the one-pass kernel (4, 6, 8, 12 or 16)
    // sources
    int nsrc = 0;
    int *d_src = nullptr;  // [kMaxSrc][4] (i, j, k, component)
    void *d_src_val = nullptr;
    // step counter and monitor ring
    int64_t *d_step = nullptr;
    int64_t host_step = 0;
    double *d_ring = nullptr;
    int64_t ring_rows = 0, mon_read = 0;
    bool mon_on = false;
    int64_t mon_c = -1, mon_ridx = -1, mon_vidx = -1;
    const void *mon_r = nullptr, *mon_v = nullptr, *mon_vcb = nullptr;
    int64_t mon_rc = -1, mon_vc = -1;
    unsigned long long *d_diag = nullptr;
    unsigned long long *d_diag_part = nullptr;  // [kDiagWaves][2] wavefront pairs of the lattice diagnostics
    float *d_slice = nullptr;
    // near-field DFT of one plane
    bool dft_on = false;
    int dft_comp = 0;
    DftArgs dft{};
    double2 *d_dft_w = nullptr, *d_dft_acc = nullptr;
    // argument blocks of the current plan (batched launches read these)
    YeeArgs<double> yee64{};
    YeeArgs<float> yee32{};
    StepEndArgs se{};
    // batch membership: the batch's stream runs this context's steps
    phpfdtd_batch *batch = nullptr;
    hipEvent_t ev = nullptr;  // marks this stream's latest work for the batch stream
    // plan and graphs
    std::vector<Kernel> plan;
    hipGraphExec_t g1 = nullptr, g16 = nullptr;
    bool dirty = true;
    bool use_graph = true;

    ~phpfdtd_ctx() {
        if (!stream && allocs.empty()) return;  // failed before touching the device
        (void)hipSetDevice(device);
        if (stream) (void)hipStreamSynchronize(stream);
        if (g1) (void)hipGraphExecDestroy(g1);
        if (g16) (void)hipGraphExecDestroy(g16);
        if (ev) (void)hipEventDestroy(ev);
        for (void *p : allocs) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }

    int alloc(void **p, size_t n) {
        *p = nullptr;
        if (n == 0) n = 8;
        hipError_t e = hipMalloc(p, n);
        if (e != hipSuccess) {
            *p = nullptr;
            (void)hipGetLastError();
            return fail(PHPFDTD_ENOMEM_HBM, "hipMalloc(%zu) failed: %s (context holds %zu bytes)", n,
                        hipGetErrorString(e), bytes);
        }
        allocs.push_back(*p);
        bytes += n;
        e = hipMemsetAsync(*p, 0, n, stream);
        if (e != hipSuccess) return fail(PHPFDTD_EHIP, "hipMemsetAsync: %s", hipGetErrorString(e));
        return 0;
    }

    template <typename T>
    int upload(void **p, const std::vector<T> &v) {
        int rc = alloc(p, v.size() * sizeof(T));
        if (rc) return rc;
        if (!v.empty()) HIPCHK(hipMemcpyAsync(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, stream));
        return 0;
    }

    int upload_real(void **p, const std::vector<double> &v) {
        if (prec == PHPFDTD_FP64) return upload(p, v);
        std::vector<float> f(v.begin(), v.end());
        return upload(p, f);
    }

    size_t fidx(int i, int j, int k) const { return ((size_t)k * ny + j) * nxp + i; }

    // component q (0..8) of the current E/H set (after host_step steps)
    void *cur(int q) const { return F[host_step & 1][q]; }

    // free one allocation made by alloc (after the stream is idle)
    void release(void *p) {
        if (!p) return;
        auto it = std::find(allocs.begin(), allocs.end(), p);
        if (it == allocs.end()) return;
        allocs.erase(it);
        (void)hipFree(p);
    }
};

static Layout make_layout(const phpfdtd_grid &g, const std::vector<int64_t> &nslot,
                          const std::vector<int> &natom, const std::vector<int> &ade) {
    const bool mfma = g.precision != PHPFDTD_FP64;
    const bool pk = g.precision == PHPFDTD_FP32;  // compensated FP32 lattice state, 5 bytes per value
    Layout l;
    const size_t rs = mfma ? 4 : 8;
    l.nxp = (g.nx + 31) / 32 * 32;
    l.ncellp = (size_t)l.nxp * g.ny * g.nz;
    l.field = l.ncellp * rs;
    const size_t L2 = 2 * (size_t)g.pml_thickness;
    l.psi_x = L2 * g.ny * g.nz * rs;
    l.psi_y = L2 * l.nxp * g.nz * rs;
    l.psi_z = L2 * l.nxp * g.ny * rs;
    l.medium = l.ncellp;
    l.small = kMaxMat * sizeof(MatCoef<double>) + 3 * ((size_t)g.nx + g.ny + g.nz) * (5 * rs + 4) +
              ((size_t)g.ny + g.nz) * sizeof(AxRow<double>) +
              (size_t)g.tsteps * kMaxSrc * rs + 1024 + 64 * 1024 + (size_t)kDiagWaves * 16;
    for (size_t m = 0; m < nslot.size(); ++m) {
        const size_t n = nslot[m];
        if (!n) continue;
        l.slots += n * 4;
        if (ade[m]) l.slots += 2 * 3 * n * rs;
        if (natom[m]) {
            const IonLayout il = ion_layout(natom[m], pk);
            l.slots += (n + 63) / 64 * (size_t)(il.tile_r() + il.tile_v()) * (pk ? 4 : 8) + 9 * natom[m] * natom[m] * 8 +
                       18 * natom[m] * 8;
            if (mfma)
                l.slots += ((size_t)mfma_nma(natom[m]) * mfma_nk(natom[m]) * 64 + 18 * natom[m]) * sizeof(float);
        }
    }
    l.ring = (size_t)g.tsteps * 12 * 8;
    return l;
}

static int check_inputs(const phpfdtd_grid *g, const phpfdtd_material *mats, int nmat, const int32_t *medium) {
    if (!g || !mats || !medium) return fail(PHPFDTD_EINVAL, "null argument");
    std::string err;
    if (validate_grid(*g, &err)) return fail(PHPFDTD_EINVAL, "%s", err.c_str());
    if (nmat < 1 || nmat > kMaxMat) return fail(PHPFDTD_EINVAL, "nmat must be in [1, %d]", kMaxMat);
    return 0;
}

// medium (file order) -> per-material cell counts; validates ids (the
// reference never checks them, fdtd.c:3408).
static int count_medium(const phpfdtd_grid &g, const int32_t *medium, int nmat, std::vector<int64_t> *cnt) {
    cnt->assign(nmat, 0);
    const size_t n = (size_t)g.nx * g.ny * g.nz;
    for (size_t c = 0; c < n; ++c) {
        const int32_t id = medium[c];
        if (id < 0 || id >= nmat)
            return fail(PHPFDTD_EFORMAT, "medium id %d at cell %zu is outside the material table [0, %d)", id, c,
                        nmat);
        (*cnt)[id]++;
    }
    return 0;
}

// Operand tables of the precision-form lattice kernel:
//   aop: -DMM as v_mfma_f32_16x16x4f32 A operands, A[m][k] of m-tile mt and
//        k-tile kt in lane m % 16 + 16 (k % 4) at aop[(mt*nk + kt)*64 + lane];
//   qop: [QeZV | QeZM] in FP32, 9N each, [atom][E component][direction].
static void mfma_tables(const Special &sp, std::vector<float> *aop, std::vector<float> *qop) {
    const int N = sp.pre.natom, N3 = 3 * N;
    const int NK = mfma_nk(N), NA = mfma_nma(N);
    aop->assign((size_t)NA * NK * 64, 0.f);
    for (int mt = 0; mt < NA; ++mt)
        for (int kt = 0; kt < NK; ++kt)
            for (int l = 0; l < 64; ++l) {
                const int m = 16 * mt + l % 16, k = 4 * kt + l / 16;
                if (m < N3 && k < N3) (*aop)[((size_t)mt * NK + kt) * 64 + l] = (float)-sp.dmm[(size_t)m * N3 + k];
            }
    qop->assign((size_t)18 * N, 0.f);
    for (int q = 0; q < 9 * N; ++q) {
        (*qop)[q] = (float)sp.qezv[q];
        (*qop)[9 * N + q] = (float)sp.qezm[q];
    }
}

template <typename real> static YeeArgs<real> &args_of(phpfdtd_ctx *x);
template <> YeeArgs<double> &args_of<double>(phpfdtd_ctx *x) { return x->yee64; }
template <> YeeArgs<float> &args_of<float>(phpfdtd_ctx *x) { return x->yee32; }
template <typename real> static SlotArgs<real> &slot_args_of(Special &sp);
template <> SlotArgs<double> &slot_args_of<double>(Special &sp) { return sp.sa64; }
template <> SlotArgs<float> &slot_args_of<float>(Special &sp) { return sp.sa32; }

// Yee plan switches, read at creation and whenever the plan is rebuilt (so a
// context can be timed under several settings): PHPFDTD_YEE=split runs H and
// D/E as two kernels (same results) instead of the one-pass kernel;
// PHPFDTD_YEE_KCHUNK sets the z planes per tile and PHPFDTD_YEE_ROWS the y rows
// per work-group of the one-pass kernel (defaults measured on MI355X,
// 1024x1024x640).  A batch member runs the batch's split plan instead.
static void plan_env(phpfdtd_ctx *x) {
    if (x->batch) return;
    const char *yp = getenv("PHPFDTD_YEE");
    x->fused = !(yp && strcmp(yp, "split") == 0);
    x->yee_kchunk = x->prec == PHPFDTD_FP32 ? (x->fused ? 12 : 4) : 16;
    if (const char *e = getenv("PHPFDTD_YEE_KCHUNK")) {
        const int v = atoi(e);
        if (v >= 1) x->yee_kchunk = v;
    }
    // Measured on the 1024x1024x640 case.  FP32 (E-form, 127 VGPRs, 4
    // wavefronts per SIMD): 12-plane chunks 7.62-7.66 ms against 7.73-7.74
    // for 16 (32: 7.91, 64: 8.29); 8 rows (two work-groups per CU) 7.7 ms
    // against 8.1 for 16 rows (one work-group of 16 wavefronts), 9.3 for 4
    // rows.  FP64: 16-plane chunks; 8 rows 21.70 ms against 22.44 (4) and
    // 23.25 (12: LDS-bound to one work-group per CU) before the band order.
    x->yee_rows = 8;
    if (const char *e = getenv("PHPFDTD_YEE_ROWS")) {
        const int v = atoi(e);
        if (v == 4 || v == 6 || v == 8 || v == 12 || v == 16) x->yee_rows = v;
    }
}

// derive_d over the box [b0, b1) x [b2, b3) x [b4, b5) of cells, reading E
// set (step + ahead) & 1.
static DeriveArgs derive_args(const phpfdtd_ctx *x, const int b[6], int ahead) {
    DeriveArgs d{};
    for (int set = 0; set < 2; ++set)
        for (int q = 0; q < 3; ++q) d.E[set][q] = x->F[set][q];
    for (int q = 0; q < 3; ++q) d.D[q] = x->F[0][3 + q];
    d.medium = x->d_medium;
    d.dm = x->d_dm;
    d.derive = x->d_derive;
    d.nxp = x->nxp;
    d.ny = x->ny;
    d.i0 = b[0]; d.i1 = b[1]; d.j0 = b[2]; d.j1 = b[3]; d.k0 = b[4]; d.k1 = b[5];
    d.ahead = ahead;
    d.step = x->d_step;
    return d;
}

// Between steps: D of the E-form plain cells of a box, written to the D
// arrays on the context's stream before a reader copies them (D-form: no-op).
static void derive_box(phpfdtd_ctx *x, int i0, int i1, int j0, int j1, int k0, int k1) {
    if (!x->eform) return;
    const int b[6] = {i0, i1, j0, j1, k0, k1};
    launch_derive<float>(derive_args(x, b, 0), x->stream);
}

template <typename real>
static void build_plan_t(phpfdtd_ctx *x) {
    plan_env(x);
    x->plan.clear();
    const size_t w = sizeof(real);
    YeeArgs<real> a{};
    for (int set = 0; set < 2; ++set) {
        for (int q = 0; q < 3; ++q) {
            a.E[set][q] = (real *)x->F[set][q];
            a.H[set][q] = (real *)x->F[set][6 + q];
        }
        const int hpsi[6] = {0, 1, 4, 5, 8, 9};  // Hxy Hxz Hyx Hyz Hzx Hzy
        for (int q = 0; q < 6; ++q) a.pH[set][q] = (real *)x->psi[set][hpsi[q]];
    }
    a.Dx = (real *)x->F[0][3]; a.Dy = (real *)x->F[0][4]; a.Dz = (real *)x->F[0][5];
    a.pDxy = (real *)x->psi[0][2]; a.pDxz = (real *)x->psi[0][3];
    a.pDyx = (real *)x->psi[0][6]; a.pDyz = (real *)x->psi[0][7];
    a.pDzx = (real *)x->psi[0][10]; a.pDzy = (real *)x->psi[0][11];
    a.medium = x->d_medium;
    a.mats = (const MatCoef<real> *)x->d_mats;
    a.ax.b = (const real *)x->d_axis[0][0];
    a.ax.kh = (const real *)x->d_axis[0][1];
    a.ax.kd = (const real *)x->d_axis[0][2];
    a.ax.ch = (const real *)x->d_axis[0][3];
    a.ax.cd = (const real *)x->d_axis[0][4];
    a.ax.slot = x->d_slot[0];
    a.ay = (const AxRow<real> *)x->d_axrow[1];
    a.az = (const AxRow<real> *)x->d_axrow[2];
    a.nx = x->nx; a.ny = x->ny; a.nz = x->nz; a.nxp = x->nxp; a.L = x->L;
    a.nmat = x->nmat;
    // a work-group is `rows` wavefronts (4 in the split plan), each one y row
    // x 64 lanes x one 16-byte vector of x cells, marching over a short chunk
    // of z planes (streams of many short-lived wavefronts run closest to the
    // HBM peak)
    const int vec = YeeVec<real>::value;
    a.nbx = (x->nx + 64 * vec - 1) / (64 * vec);
    a.rows = x->fused ? x->yee_rows : kBlockY;
    a.nby = (x->ny + a.rows - 1) / a.rows;
    const int kc = std::min(x->yee_kchunk, x->nz);
    a.kchunk = kc;
    a.kbase = 0;
    a.kend = x->nz;
    a.nbz = (x->nz + kc - 1) / kc;
    a.nblocks = a.nbx * a.nby * a.nbz;
    a.nblocks_pad = (a.nblocks + 8 * a.nbx - 1) / (8 * a.nbx) * (8 * a.nbx);  // whole groups of band_tile
    a.nsrc = x->nsrc;
    a.src = x->d_src;
    a.src_val = (const real *)x->d_src_val;
    a.step = x->d_step;
    a.dmask = x->dmask;
    args_of<real>(x) = a;

    // algorithmic bytes (compulsory traffic of each kernel's own streams)
    const double nx = x->nx, ny = x->ny, nz = x->nz, L = x->L;
    const double cells = nx * ny * nz;
    const double hcells = (nx - 1) * (ny - 1) * nz;
    const double icells = (nx - 2) * (ny - 2) * nz;
    auto pml_inc = [&](double ex, double ey, double ez) {  // (cell, dir) incidences in the region
        const double lx = std::min(2 * L, ex), ly = std::min(2 * L, ey), lz = std::min(2 * L, nz);
        return lx * ey * ez + ly * ex * ez + lz * ex * ey;
    };
    double special_cells = 0;
    for (const auto &sp : x->specials) special_cells += (double)sp.cells.size();
    const double lossy_int = (double)x->lossy_interior;
    const double psi_h = pml_inc(nx - 1, ny - 1, nz) * 4 * w, psi_d = pml_inc(nx - 2, ny - 2, nz) * 4 * w;
    // D read and written: every cell (D-form), cells that keep D (E-form)
    const double dcells = x->eform ? (double)x->dstate_cells : cells;
    if (x->fused) {
        // E read once (3w), E of non-slot cells written (3w), H read and
        // written (6w), D read and written (6w), material id, psi of both halves
        Kernel ky;
        ky.name = "yee_update[r" + std::to_string(a.rows) + "]";  // rows per work-group
        ky.stage = K_YEE;
        ky.bytes = cells * (3 * w + 1) + dcells * 6 * w + (cells - special_cells) * 3 * w + hcells * 6 * w + psi_h +
                   psi_d;
        ky.flops = hcells * 18 + icells * 18 + cells * 15;
        ky.fn = [a](hipStream_t s) { launch_yee<real>(a, s); };
        x->plan.push_back(ky);
    } else {
        Kernel kh;
        kh.name = "h_update";
        kh.stage = K_YEE;
        kh.bytes = hcells * 9 * w + psi_h;
        kh.flops = hcells * 18;
        kh.fn = [a](hipStream_t s) { launch_h<real>(a, s); };
        x->plan.push_back(kh);
        Kernel kd;
        kd.name = "de_update";
        kd.stage = K_YEE;
        // E-form: E^n of every cell is read (it is advanced); D-form: only lossy cells' E
        kd.bytes = cells * 1 + dcells * 6 * w + (cells - special_cells) * 3 * w + icells * 3 * w +
                   (x->eform ? cells : lossy_int) * 3 * w + psi_d;
        kd.flops = icells * 18 + cells * 15;
        kd.fn = [a](hipStream_t s) { launch_de<real>(a, s); };
        x->plan.push_back(kd);
    }

    for (auto &sp : x->specials) {
        SlotArgs<real> sa{};
        for (int set = 0; set < 2; ++set)
            for (int q = 0; q < 3; ++q) sa.E[set][q] = a.E[set][q];
        sa.Dx = a.Dx; sa.Dy = a.Dy; sa.Dz = a.Dz;
        sa.step = x->d_step;
        sa.cells = sp.d_cells;
        sa.nslot = (int64_t)sp.cells.size();
        for (int e = 0; e < 9; ++e) {
            sa.R[e] = (real)sp.pre.R[e];
            sa.A[e] = (real)sp.pre.A[e];
            sa.Gpi[e] = (real)sp.pre.Gpi[e];
            sa.Gm[e] = (real)sp.pre.Gm[e];
            sa.W[e] = (real)sp.pre.W[e];
            sa.Ga[e] = (real)sp.pre.Ga[e];
            sa.Gb[e] = (real)sp.pre.Gb[e];
        }
        sa.ade = sp.pre.ade ? 1 : 0;
        sa.P = (real *)sp.d_P;
        sa.Po = (real *)sp.d_Po;
        sa.natom = sp.pre.natom;
        sa.r = sp.d_r; sa.v = sp.d_v;
        sa.pk = x->pk ? 1 : 0;
        sa.lay = sp.lay;
        sa.dmm = sp.d_dmm; sa.qezv = sp.d_qezv; sa.qezm = sp.d_qezm;
        sa.aop = sp.d_aop; sa.qop = sp.d_qop;
        sa.dt = x->g.dt;
        sa.half_dt2 = 0.5 * x->g.dt * x->g.dt;
        const double n = (double)sp.cells.size();
        Kernel k;
        char nm[48];
        if (sp.pre.natom > 0) {
            const double N3 = 3.0 * sp.pre.natom;
            snprintf(nm, sizeof(nm), "ion_update[m%d,N%d]", sp.id, sp.pre.natom);
            // r, v read and written: 8 bytes per value, 4.5 for compensated FP32
            k.bytes = n * (6 * w + 4 + 4 * N3 * (x->pk ? 4.5 : 8) + (sp.pre.ade ? 12 * w : 0));
            k.flops = n * (2 * N3 * N3 + 12 * N3 + 40);
            k.stage = K_ION;
            k.fn = [sa](hipStream_t s) { launch_ion<real>(sa, s); };
        } else {
            snprintf(nm, sizeof(nm), "ade_update[m%d]", sp.id);
            k.bytes = n * (21 * w + 4);
            k.flops = n * 80;
            k.fn = [sa](hipStream_t s) { launch_ade<real>(sa, s); };
        }
        k.name = nm;
        x->plan.push_back(k);
        slot_args_of<real>(sp) = sa;
    }
    if (x->dft_on && x->eform && x->dft_comp >= PHPFDTD_DX && x->dft_comp <= PHPFDTD_DZ) {
        const DftArgs &da = x->dft;
        int b[6] = {0, x->nx, 0, x->ny, 0, x->nz};
        const int ax = da.axis == 'x' ? 0 : da.axis == 'y' ? 1 : 2;
        b[2 * ax] = da.pos;
        b[2 * ax + 1] = std::min(da.pos + 1, ax == 0 ? x->nx : ax == 1 ? x->ny : x->nz);
        const DeriveArgs dv = derive_args(x, b, 1);  // inside the step: the E set just written
        Kernel kv;
        kv.name = "derive_d";
        kv.bytes = (double)da.nplane * (3 * w + 1);
        kv.stage = K_PRE;
        kv.fn = [dv](hipStream_t s) { launch_derive<float>(dv, s); };
        x->plan.push_back(kv);
    }
    if (x->dft_on) {
        DftArgs da = x->dft;
        Kernel kf;
        kf.name = "dft_accum";
        kf.bytes = (double)da.nplane * (w + 32);
        kf.flops = (double)da.nplane * 4;
        // D and H are final before the next H update, E only after the lattice kernels
        kf.stage = (x->dft_comp >= PHPFDTD_DX && x->dft_comp <= PHPFDTD_HZ) ? K_PRE : K_POST;
        kf.fn = [da](hipStream_t s) { launch_dft<real>(da, s); };
        x->plan.push_back(kf);
    }
    StepEndArgs se{};
    se.step = x->d_step;
    se.ring = x->d_ring;
    se.ring_rows = x->ring_rows;
    se.mon_on = x->mon_on ? 1 : 0;
    se.mon_c = x->mon_c;
    for (int set = 0; set < 2; ++set)
        for (int q = 0; q < 9; ++q) se.F[set][q] = x->F[set][q];
    se.mon_r = x->mon_r;
    se.mon_v = x->mon_v;
    se.mon_rc = x->mon_rc;
    se.mon_vc = x->mon_vc;
    se.mon_vcb = x->mon_vcb;
    se.mon_ridx = x->mon_ridx;
    se.mon_vidx = x->mon_vidx;
    se.dt = x->g.dt;
    se.part = 0;
    se.mon_dfe = 0;
    if (x->eform && x->mon_c >= 0 && x->derive[x->medium[x->mon_c]]) {
        se.mon_dfe = 1;
        memcpy(se.mon_dm, &x->dm[9 * (size_t)x->medium[x->mon_c]], sizeof(se.mon_dm));
    }
    x->se = se;
    Kernel ke;
    ke.name = "step_end";
    ke.stage = K_POST;
    ke.fn = [se](hipStream_t s) { launch_step_end<real>(se, s); };
    x->plan.push_back(ke);
    // issue order of one step
    std::stable_sort(x->plan.begin(), x->plan.end(),
                     [](const Kernel &p, const Kernel &q) { return p.stage < q.stage; });
}

static void build_plan(phpfdtd_ctx *x) {
    if (x->prec == PHPFDTD_FP64) build_plan_t<double>(x);
    else build_plan_t<float>(x);
}

static void touch(phpfdtd_ctx *x);

static int rebuild(phpfdtd_ctx *x) {
    if (x->g1) { (void)hipGraphExecDestroy(x->g1); x->g1 = nullptr; }
    if (x->g16) { (void)hipGraphExecDestroy(x->g16); x->g16 = nullptr; }
    build_plan(x);
    x->dirty = false;
    if (!x->use_graph) return 0;
    for (int which = 0; which < 2; ++which) {
        const int n = which == 0 ? 1 : 16;
        hipGraph_t graph = nullptr;
        HIPCHK(hipStreamBeginCapture(x->stream, hipStreamCaptureModeThreadLocal));
        for (int t = 0; t < n; ++t)
            for (auto &k : x->plan) k.fn(x->stream);
        HIPCHK(hipStreamEndCapture(x->stream, &graph));
        hipGraphExec_t ex = nullptr;
        hipError_t e = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return fail(PHPFDTD_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
        (which == 0 ? x->g1 : x->g16) = ex;
    }
    return 0;
}

template <typename real>
static std::vector<real> cast_vec(const std::vector<double> &v) {
    return std::vector<real>(v.begin(), v.end());
}

// Update form of each material.  The FP32 form runs plain cells in E-form
// (E^{n+1} = E^n + M . dD, D not stored); slot materials keep D, and so does
// a plain material whose M is singular (|det M| below 1e-12 of the product
// of its row norms), since its D could not be recovered from E.
static void eform_tables(phpfdtd_ctx *x, const std::vector<int64_t> &cnt) {
    x->eform = x->prec == PHPFDTD_FP32;
    x->dstate.assign(x->nmat, 1);
    x->derive.assign(x->nmat, 0);
    x->dm.assign(9 * (size_t)x->nmat, 0.0);
    x->dmask = 0;
    x->dstate_cells = 0;
    for (int m = 0; m < x->nmat; ++m) {
        const double *M = x->pre[m].M;
        bool keep = !x->eform || x->pre[m].special;
        if (!keep) {
            const double det = M[0] * (M[4] * M[8] - M[5] * M[7]) - M[1] * (M[3] * M[8] - M[5] * M[6]) +
                               M[2] * (M[3] * M[7] - M[4] * M[6]);
            double nrm = 1.0;
            for (int r = 0; r < 3; ++r)
                nrm *= sqrt(M[3 * r] * M[3 * r] + M[3 * r + 1] * M[3 * r + 1] + M[3 * r + 2] * M[3 * r + 2]);
            keep = !(fabs(det) > 1e-12 * nrm) || inverse3(M, &x->dm[9 * (size_t)m]) != 0;
        }
        x->dstate[m] = keep ? 1 : 0;
        x->derive[m] = keep ? 0 : 1;
        if (keep) {
            x->dmask |= 1ull << m;
            x->dstate_cells += cnt[m];
        }
    }
}

template <typename real>
static int upload_mats(phpfdtd_ctx *x) {
    std::vector<MatCoef<real>> mc(x->nmat);
    for (int m = 0; m < x->nmat; ++m) {
        memset(&mc[m], 0, sizeof(mc[m]));
        for (int e = 0; e < 9; ++e) {
            mc[m].M[e] = (real)x->pre[m].M[e];
            mc[m].Lk[e] = (real)x->pre[m].Lk[e];
        }
        mc[m].flags = (x->pre[m].loss ? MF_LOSS : 0) | (x->pre[m].special ? MF_SPECIAL : 0) |
                      (x->dstate[m] ? MF_DSTATE : 0);
    }
    return x->upload(&x->d_mats, mc);
}

// Coefficient rows of a y or z axis, packed (AxRow).
template <typename real>
static int upload_rows(phpfdtd_ctx *x, const AxisHost &ax, void **out) {
    std::vector<AxRow<real>> rows(ax.kh.size());
    for (size_t i = 0; i < rows.size(); ++i) {
        memset(&rows[i], 0, sizeof(rows[i]));
        rows[i].kh = (real)ax.kh[i];
        rows[i].kd = (real)ax.kd[i];
        rows[i].b = (real)ax.b[i];
        rows[i].ch = (real)ax.ch[i];
        rows[i].cd = (real)ax.cd[i];
        rows[i].p = ax.slot[i];
    }
    return x->upload(out, rows);
}

// Every int entry point runs its body through guarded(): a C++ exception
// (a host allocation failure, say) becomes an error code instead of crossing
// the C ABI and ending the caller's process.
template <typename F>
static int guarded(F &&f) {
    try {
        return f();
    } catch (const std::bad_alloc &) {
        return fail(PHPFDTD_ENOMEM_HOST, "host memory allocation failed");
    } catch (const std::exception &e) {
        return fail(PHPFDTD_EINVAL, "internal error: %s", e.what());
    } catch (...) {
        return fail(PHPFDTD_EINVAL, "internal error");
    }
}

extern "C" {

int phpfdtd_abi_version(void) { return PHPFDTD_ABI_VERSION; }

double phpfdtd_source_value(const phpfdtd_source *src, double t) { return src ? source_value(*src, t) : 0.0; }

const char *phpfdtd_last_error(void) { return g_err; }

static int footprint_bytes_impl(const phpfdtd_grid *g, const phpfdtd_material *mats, int nmat, const int32_t *medium,
                            size_t *bytes) {
    int rc = check_inputs(g, mats, nmat, medium);
    if (rc) return rc;
    std::vector<int64_t> cnt;
    if ((rc = count_medium(*g, medium, nmat, &cnt))) return rc;
    std::vector<int> natom(nmat), ade(nmat);
    for (int m = 0; m < nmat; ++m) {
        natom[m] = mats[m].natom;
        ade[m] = mats[m].use_ade;
        if (!natom[m] && !ade[m]) cnt[m] = 0;
    }
    *bytes = make_layout(*g, cnt, natom, ade).total();
    return 0;
}

int phpfdtd_footprint_bytes(const phpfdtd_grid *g, const phpfdtd_material *mats, int nmat, const int32_t *medium,
                            size_t *bytes) {
    return guarded([&] { return footprint_bytes_impl(g, mats, nmat, medium, bytes); });
}

static int create_impl(const phpfdtd_grid *g, const phpfdtd_material *mats, int nmat, const int32_t *medium,
                   phpfdtd_ctx **out) {
    if (!out) return fail(PHPFDTD_EINVAL, "null out");
    *out = nullptr;
    int rc = check_inputs(g, mats, nmat, medium);
    if (rc) return rc;
    std::vector<int64_t> cnt;
    if ((rc = count_medium(*g, medium, nmat, &cnt))) return rc;

    phpfdtd_ctx *x = new phpfdtd_ctx();
    std::unique_ptr<phpfdtd_ctx> guard(x);
    x->g = *g;
    x->prec = g->precision == PHPFDTD_FP64 ? PHPFDTD_FP64 : PHPFDTD_FP32;
    x->pk = g->precision == PHPFDTD_FP32;
    x->rs = x->prec == PHPFDTD_FP32 ? 4 : 8;
    x->nx = g->nx; x->ny = g->ny; x->nz = g->nz; x->L = g->pml_thickness;
    x->nmat = nmat;
    const char *ng = getenv("PHPFDTD_NO_GRAPH");
    x->use_graph = !(ng && ng[0] == '1');
    plan_env(x);
    x->pre.resize(nmat);
    for (int m = 0; m < nmat; ++m) {
        std::string err;
        if ((rc = material_coefs(*g, mats[m], &x->pre[m], &err)))
            return fail(rc, "material %d: %s", m, err.c_str());
    }
    HIPCHK(hipGetDevice(&x->device));
    HIPCHK(hipStreamCreateWithFlags(&x->stream, hipStreamNonBlocking));

    std::vector<int> natom(nmat), ade(nmat);
    for (int m = 0; m < nmat; ++m) {
        natom[m] = mats[m].natom;
        ade[m] = mats[m].use_ade;
    }
    std::vector<int64_t> scnt(cnt);
    for (int m = 0; m < nmat; ++m)
        if (!x->pre[m].special) scnt[m] = 0;
    const Layout lay = make_layout(*g, scnt, natom, ade);
    x->nxp = lay.nxp;
    x->ncellp = lay.ncellp;
    x->plane = (size_t)x->nxp * x->ny;

    // E and H in two sets, D in one; H psi slabs in two sets, D psi in one
    for (int set = 0; set < 2; ++set)
        for (int q = 0; q < 9; ++q) {
            const bool d = q >= 3 && q < 6;
            if (set == 1 && d) x->F[1][q] = x->F[0][q];
            else if ((rc = x->alloc(&x->F[set][q], lay.field))) return rc;
        }
    if (x->L > 0)
        for (int set = 0; set < 2; ++set)
            for (int q = 0; q < 12; ++q) {
                const bool d = (q & 3) >= 2;  // Dxy Dxz | Dyx Dyz | Dzx Dzy
                if (set == 1 && d) x->psi[1][q] = x->psi[0][q];
                else if ((rc = x->alloc(&x->psi[set][q], q < 4 ? lay.psi_x : q < 8 ? lay.psi_y : lay.psi_z)))
                    return rc;
            }
    // medium: file order [nx][ny][nz] -> x fastest; padding cells get id 0
    x->medium.assign(x->ncellp, 0);
    for (int i = 0; i < x->nx; ++i)
        for (int j = 0; j < x->ny; ++j) {
            const int32_t *col = medium + ((size_t)i * x->ny + j) * x->nz;
            for (int k = 0; k < x->nz; ++k) x->medium[x->fidx(i, j, k)] = (uint8_t)col[k];
        }
    if ((rc = x->upload((void **)&x->d_medium, x->medium))) return rc;
    eform_tables(x, cnt);
    if (x->prec == PHPFDTD_FP64) rc = upload_mats<double>(x);
    else rc = upload_mats<float>(x);
    if (rc) return rc;
    if (x->eform) {
        if ((rc = x->upload((void **)&x->d_dm, x->dm))) return rc;
        if ((rc = x->upload((void **)&x->d_derive, x->derive))) return rc;
    }
    const int ns[3] = {x->nx, x->ny, x->nz};
    const double ds[3] = {g->dx, g->dy, g->dz};
    for (int d = 0; d < 3; ++d) {
        AxisHost ax = axis_tables(ns[d], ds[d], *g);
        if (d == 0) {  // x tables are read as 16-byte vectors over the padded row
            ax.b.resize(x->nxp, 0.0);
            ax.kh.resize(x->nxp, 0.0);
            ax.kd.resize(x->nxp, 0.0);
            ax.ch.resize(x->nxp, 0.0);
            ax.cd.resize(x->nxp, 0.0);
            ax.slot.resize(x->nxp, -1);
        }
        const std::vector<double> *tabs[5] = {&ax.b, &ax.kh, &ax.kd, &ax.ch, &ax.cd};
        for (int t = 0; t < 5; ++t)
            if ((rc = x->upload_real(&x->d_axis[d][t], *tabs[t]))) return rc;
        if ((rc = x->upload((void **)&x->d_slot[d], ax.slot))) return rc;
        if (d > 0) {  // y rows and z planes, packed
            if (x->prec == PHPFDTD_FP64) rc = upload_rows<double>(x, ax, &x->d_axrow[d]);
            else rc = upload_rows<float>(x, ax, &x->d_axrow[d]);
            if (rc) return rc;
        }
    }
    // slot lists of ADE / lattice materials, x-fastest order
    for (int m = 0; m < kMaxMat; ++m) x->special_of[m] = -1;
    for (int m = 0; m < nmat; ++m) {
        if (!x->pre[m].special || cnt[m] == 0) continue;
        Special sp;
        sp.id = m;
        sp.pre = x->pre[m];
        sp.cells.reserve(cnt[m]);
        const int n = mats[m].natom;
        if (n > 0) {
            sp.qezv.assign(mats[m].qezv, mats[m].qezv + 9 * n);
            sp.qezm.assign(mats[m].qezm, mats[m].qezm + 9 * n);
            sp.dmm.assign(mats[m].dmm, mats[m].dmm + 9 * n * n);
        }
        x->special_of[m] = (int)x->specials.size();
        x->specials.push_back(std::move(sp));
    }
    for (size_t c = 0; c < x->ncellp; ++c) {
        const int i = (int)(c % x->nxp);
        if (i >= x->nx) continue;
        const int so = x->special_of[x->medium[c]];
        if (so >= 0) x->specials[so].cells.push_back((uint32_t)c);
        const int j = (int)((c / x->nxp) % x->ny);
        if (x->pre[x->medium[c]].loss && i >= 1 && i <= x->nx - 2 && j >= 1 && j <= x->ny - 2) x->lossy_interior++;
    }
    for (auto &sp : x->specials) {
        const size_t n = sp.cells.size();
        if ((rc = x->upload((void **)&sp.d_cells, sp.cells))) return rc;
        if (sp.pre.ade) {
            if ((rc = x->alloc(&sp.d_P, 3 * n * x->rs))) return rc;
            if ((rc = x->alloc(&sp.d_Po, 3 * n * x->rs))) return rc;
        }
        if (sp.pre.natom > 0) {
            sp.lay = ion_layout(sp.pre.natom, x->pk);
            const size_t ntile = (n + 63) / 64;  // 64-slot tiles
            const size_t es = x->pk ? 4 : 8;  // bytes per element (IonLayout)
            if ((rc = x->alloc(&sp.d_r, ntile * sp.lay.tile_r() * es))) return rc;
            if ((rc = x->alloc(&sp.d_v, ntile * sp.lay.tile_v() * es))) return rc;
            if ((rc = x->upload((void **)&sp.d_dmm, sp.dmm))) return rc;
            if ((rc = x->upload((void **)&sp.d_qezv, sp.qezv))) return rc;
            if ((rc = x->upload((void **)&sp.d_qezm, sp.qezm))) return rc;
            if (x->prec == PHPFDTD_FP32) {
                std::vector<float> aop, qop;
                mfma_tables(sp, &aop, &qop);
                if ((rc = x->upload((void **)&sp.d_aop, aop))) return rc;
                if ((rc = x->upload((void **)&sp.d_qop, qop))) return rc;
            }
        }
    }
    if ((rc = x->alloc((void **)&x->d_step, 64))) return rc;
    x->ring_rows = g->tsteps;
    if ((rc = x->alloc((void **)&x->d_ring, (size_t)std::max(1, g->tsteps) * 12 * 8))) return rc;
    if ((rc = x->alloc((void **)&x->d_diag, 64))) return rc;
    if ((rc = x->alloc((void **)&x->d_diag_part, (size_t)kDiagWaves * 2 * sizeof(unsigned long long)))) return rc;
    if ((rc = x->alloc(&x->d_src_val, x->rs))) return rc;
    if ((rc = x->alloc((void **)&x->d_src, 4 * kMaxSrc * sizeof(int)))) return rc;
    {
        const size_t mx = std::max({(size_t)x->nx * x->ny, (size_t)x->ny * x->nz, (size_t)x->nx * x->nz});
        if ((rc = x->alloc((void **)&x->d_slice, mx * sizeof(float)))) return rc;
    }
    HIPCHK(hipStreamSynchronize(x->stream));
    *out = guard.release();
    return 0;
}

int phpfdtd_create(const phpfdtd_grid *g, const phpfdtd_material *mats, int nmat, const int32_t *medium,
                   phpfdtd_ctx **out) {
    return guarded([&] { return create_impl(g, mats, nmat, medium, out); });
}

void phpfdtd_destroy(phpfdtd_ctx *x) {
    if (!x) return;
    if (x->batch) phpfdtd_batch_remove(x->batch, x);
    delete x;
}

int64_t phpfdtd_step(const phpfdtd_ctx *x) { return x ? x->host_step : -1; }

static int set_sources_impl(phpfdtd_ctx *x, int n, const phpfdtd_source *src) {
    if (!x || (n > 0 && !src) || n < 0) return fail(PHPFDTD_EINVAL, "bad arguments");
    HIPCHK(hipSetDevice(x->device));
    // keep the sources that can ever inject (fdtd.c:3375-3393 skips the rest)
    std::vector<phpfdtd_source> act;
    for (int s = 0; s < n; ++s) {
        const phpfdtd_source &q = src[s];
        const bool comp_ok = q.component == 'x' || q.component == 'y' || q.component == 'z';
        if (q.x < 0 || q.x >= x->nx || q.y < 0 || q.y >= x->ny || q.z < 0 || q.z >= x->nz || !comp_ok) continue;
        act.push_back(q);
    }
    if ((int)act.size() > kMaxSrc) return fail(PHPFDTD_EINVAL, "at most %d active sources", kMaxSrc);
    const int na = (int)act.size();
    const int T = std::max(1, x->g.tsteps);
    std::vector<double> val((size_t)T * std::max(1, na), 0.0);
    for (int t = 0; t < x->g.tsteps; ++t)
        for (int s = 0; s < na; ++s) val[(size_t)t * na + s] = source_value(act[s], t * x->g.dt) * x->g.dt;
    HIPCHK(hipStreamSynchronize(x->stream));
    void *old = x->d_src_val;
    int rc = x->upload_real(&x->d_src_val, val);
    if (rc) return rc;
    auto it = std::find(x->allocs.begin(), x->allocs.end(), old);
    if (it != x->allocs.end()) {
        x->allocs.erase(it);
        (void)hipFree(old);
    }
    std::vector<int> pos(4 * kMaxSrc, 0);
    for (int s = 0; s < na; ++s) {
        pos[4 * s] = act[s].x;
        pos[4 * s + 1] = act[s].y;
        pos[4 * s + 2] = act[s].z;
        pos[4 * s + 3] = act[s].component - 'x';
    }
    HIPCHK(hipMemcpyAsync(x->d_src, pos.data(), pos.size() * sizeof(int), hipMemcpyHostToDevice, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    x->nsrc = na;
    touch(x);
    return 0;
}

int phpfdtd_set_sources(phpfdtd_ctx *x, int n, const phpfdtd_source *src) {
    return guarded([&] { return set_sources_impl(x, n, src); });
}

static int find_slot(const phpfdtd_ctx *x, int sp, size_t c, int64_t *slot) {
    const auto &v = x->specials[sp].cells;
    auto it = std::lower_bound(v.begin(), v.end(), (uint32_t)c);
    if (it == v.end() || *it != (uint32_t)c) return -1;
    *slot = it - v.begin();
    return 0;
}

static int set_monitor_impl(phpfdtd_ctx *x, int mx, int my, int mz, int atom, int dir, int ref_compat) {
    if (!x) return fail(PHPFDTD_EINVAL, "null ctx");
    if (mx < 0 || mx >= x->nx || my < 0 || my >= x->ny || mz < 0 || mz >= x->nz || dir < 0 || dir > 2 || atom < 0)
        return fail(PHPFDTD_EINVAL, "monitor point (%d,%d,%d) atom %d dir %d out of range", mx, my, mz, atom, dir);
    const int zf = ref_compat ? mz + 1 : mz;
    x->mon_on = true;
    x->mon_c = zf < x->nz ? (int64_t)x->fidx(mx, my, zf) : -1;
    x->mon_ridx = x->mon_vidx = -1;
    x->mon_r = x->mon_v = x->mon_vcb = nullptr;
    x->mon_rc = x->mon_vc = -1;
    const int m = x->medium[x->fidx(mx, my, mz)];
    const int sp = x->special_of[m];
    if (sp >= 0 && atom < x->pre[m].natom) {
        int64_t slot = 0;
        find_slot(x, sp, x->fidx(mx, my, mz), &slot);
        x->mon_ridx = x->specials[sp].lay.r_index(slot, atom * 3 + dir);
        x->mon_vidx = x->specials[sp].lay.v_index(slot, atom * 3 + dir);
        x->mon_r = x->specials[sp].d_r;
        x->mon_v = x->specials[sp].d_v;
        if (x->pk) {
            bool in_r = false;
            x->mon_rc = x->specials[sp].lay.c_index(slot, atom * 3 + dir);
            x->mon_vc = x->specials[sp].lay.cv_index(slot, atom * 3 + dir, &in_r);
            x->mon_vcb = in_r ? x->mon_r : x->mon_v;
        }
    }
    touch(x);
    return 0;
}

int phpfdtd_set_monitor(phpfdtd_ctx *x, int mx, int my, int mz, int atom, int dir, int ref_compat) {
    return guarded([&] { return set_monitor_impl(x, mx, my, mz, atom, dir, ref_compat); });
}

// Plane of phpfdtd_read_slice / phpfdtd_set_dft: element count and the plane
// index actually read (ref_compat z reads pos+1, which may be the zero halo nz).
static int plane_of(const phpfdtd_ctx *x, char axis, int pos, int ref_compat, size_t *n, int *p) {
    *p = pos;
    if (axis == 'z') {
        if (pos < 0 || pos >= x->nz) return fail(PHPFDTD_EINVAL, "Error: No process contains slice z=%d", pos);
        *n = (size_t)x->nx * x->ny;
        if (ref_compat) *p = pos + 1;
    } else if (axis == 'x') {
        if (pos < 0 || pos >= x->nx)
            return fail(PHPFDTD_EINVAL, "Error: Invalid x slice position %d (range: 0-%d)", pos, x->nx - 1);
        *n = (size_t)x->ny * x->nz;
    } else if (axis == 'y') {
        if (pos < 0 || pos >= x->ny)
            return fail(PHPFDTD_EINVAL, "Error: Invalid y slice position %d (range: 0-%d)", pos, x->ny - 1);
        *n = (size_t)x->nx * x->nz;
    } else {
        return fail(PHPFDTD_EINVAL, "Error: Invalid slice axis '%c'. Must be 'x', 'y', or 'z'.", axis);
    }
    return 0;
}

// Bin of the near-field DFT: the arithmetic of frequency_analysis.c:362-390,
// with times t_n = (t0 + n*interval)*dt, T = t_{N-1} - t_0, fs = (N-1)/T; k
// = first argmin over i < N/2 of |i fs / N - f_hz|, -1 if none.
static int dft_bin(const phpfdtd_ctx *x, int N, int interval, double f_hz) {
    const double dt = x->g.dt;
    const double T = (double)(x->host_step + (int64_t)(N - 1) * interval) * dt - (double)x->host_step * dt;
    const double fs = 1.0 / (T / (N - 1));
    int k = -1;
    double best = 1e100;
    for (int i = 0; i < N / 2; ++i) {
        const double d = fabs(i * fs / N - f_hz);
        if (d < best) {
            best = d;
            k = i;
        }
    }
    return k;
}

// Weight n of bin k: exp(-2 pi i k n / N) / N, the angle reduced mod N
// exactly in integers.
static double2 dft_weight(int k, int n, int N) {
    const double th = 2.0 * M_PI * (double)(((int64_t)k * n) % N) / N;
    double2 w;
    w.x = cos(th) / N;
    w.y = -sin(th) / N;
    return w;
}

static int dft_check(phpfdtd_ctx *x, int comp, char axis, int pos, int ref_compat, int interval, int nsamples,
                     size_t *n, int *p) {
    if (!x || comp < 0 || comp > 9 || interval < 1 || nsamples < 2)
        return fail(PHPFDTD_EINVAL, "bad arguments (comp %d, interval %d, nsamples %d)", comp, interval, nsamples);
    return plane_of(x, axis, pos, ref_compat, n, p);
}

// Point the context's DFT at weights w and a zeroed accumulator acc (n
// plane elements) and put it in the step plan.
static void dft_install(phpfdtd_ctx *x, int comp, char axis, int p, int ref_compat, int interval, int N, size_t n,
                        const double2 *w, double2 *acc) {
    DftArgs &a = x->dft;
    a = DftArgs{};
    const int emag = comp == PHPFDTD_EMAG;
    for (int set = 0; set < 2; ++set) {
        a.F[set] = x->F[set][emag ? 0 : comp];
        a.Ex[set] = x->F[set][0];
        a.Ey[set] = x->F[set][1];
        a.Ez[set] = x->F[set][2];
    }
    a.emag = emag;
    a.ref_float = ref_compat ? 1 : 0;
    a.axis = axis;
    a.pos = p;
    a.nx = x->nx; a.ny = x->ny; a.nz = x->nz; a.nxp = x->nxp;
    a.ncell = (int64_t)x->ncellp;
    a.nplane = (int64_t)n;
    a.t0 = x->host_step;
    a.interval = interval;
    a.nsamp = N;
    a.w = w;
    a.acc = acc;
    a.step = x->d_step;
    x->dft_comp = comp;
    x->dft_on = true;
    touch(x);
}

static int set_dft_impl(phpfdtd_ctx *x, int comp, char axis, int pos, int ref_compat, int interval, int nsamples,
                    double f_hz, int *bin) {
    size_t n;
    int p;
    int rc = dft_check(x, comp, axis, pos, ref_compat, interval, nsamples, &n, &p);
    if (rc) return rc;
    HIPCHK(hipSetDevice(x->device));
    const int N = nsamples;
    const int k = dft_bin(x, N, interval, f_hz);
    if (k < 0) return fail(PHPFDTD_EINVAL, "Error finding target frequency index");
    std::vector<double2> w(N);
    for (int q = 0; q < N; ++q) w[q] = dft_weight(k, q, N);
    HIPCHK(hipStreamSynchronize(x->stream));
    x->release(x->d_dft_w);
    x->release(x->d_dft_acc);
    x->d_dft_w = x->d_dft_acc = nullptr;
    x->dft_on = false;
    if ((rc = x->upload((void **)&x->d_dft_w, w))) return rc;
    if ((rc = x->alloc((void **)&x->d_dft_acc, n * sizeof(double2)))) return rc;
    dft_install(x, comp, axis, p, ref_compat, interval, N, n, x->d_dft_w, x->d_dft_acc);
    if (bin) *bin = k;
    HIPCHK(hipStreamSynchronize(x->stream));
    return 0;
}

int phpfdtd_set_dft(phpfdtd_ctx *x, int comp, char axis, int pos, int ref_compat, int interval, int nsamples,
                    double f_hz, int *bin) {
    return guarded([&] { return set_dft_impl(x, comp, axis, pos, ref_compat, interval, nsamples, f_hz, bin); });
}

// Static DFT workspace of phpfdtd_set_dft_noalloc: kDftMaxSamples weights,
// then an accumulator of up to kDftMaxPlane plane cells (17 MiB per device,
// held until the process ends).  In host code `scratch` names a host-side
// stand-in, not device memory: its device address on each device comes from
// hipGetSymbolAddress, taken once per device on the first call there.
constexpr int kDftMaxSamples = 65536;
constexpr int64_t kDftMaxPlane = (int64_t)1 << 20;  // a 1024 x 1024 plane
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double2 scratch[kDftMaxSamples + kDftMaxPlane];
static std::mutex g_scratch_mu;
static double2 *g_scratch_dev[kMaxDevices];

// Device address of the workspace on device `dev` (current on the calling
// thread).
static int scratch_of(int dev, double2 **out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (dev < 0 || dev >= kMaxDevices) return fail(PHPFDTD_EINVAL, "device %d beyond the workspace table", dev);
    if (!g_scratch_dev[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(PHPFDTD_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
        g_scratch_dev[dev] = static_cast<double2 *>(ptr);
    }
    *out = g_scratch_dev[dev];
    return 0;
}

static int set_dft_noalloc_impl(phpfdtd_ctx *x, int comp, char axis, int pos, int ref_compat, int interval,
                                int nsamples, double f_hz, int *bin) {
    size_t n;
    int p;
    int rc = dft_check(x, comp, axis, pos, ref_compat, interval, nsamples, &n, &p);
    if (rc) return rc;
    if (nsamples > kDftMaxSamples || (int64_t)n > kDftMaxPlane)
        return fail(PHPFDTD_EINVAL,
                    "DFT of %d samples over %zu plane cells exceeds the static workspace (at most %d samples and "
                    "%lld cells); use phpfdtd_set_dft",
                    nsamples, n, kDftMaxSamples, (long long)kDftMaxPlane);
    HIPCHK(hipSetDevice(x->device));
    double2 *ws = nullptr;
    if ((rc = scratch_of(x->device, &ws))) return rc;
    const int N = nsamples;
    const int k = dft_bin(x, N, interval, f_hz);
    if (k < 0) return fail(PHPFDTD_EINVAL, "Error finding target frequency index");
    // earlier work of this context may still read the workspace or the
    // previous DFT buffers (kept: this entry point frees nothing)
    HIPCHK(hipStreamSynchronize(x->stream));
    x->dft_on = false;
    double2 *w = ws, *acc = ws + kDftMaxSamples;
    double2 chunk[256];  // weights staged through the stack, 4 KiB at a time
    for (int q0 = 0; q0 < N; q0 += 256) {
        const int m = std::min(256, N - q0);
        for (int q = 0; q < m; ++q) chunk[q] = dft_weight(k, q0 + q, N);
        HIPCHK(hipMemcpyAsync(w + q0, chunk, m * sizeof(double2), hipMemcpyHostToDevice, x->stream));
        HIPCHK(hipStreamSynchronize(x->stream));
    }
    // the accumulator starts at zero: a fill kernel, no memset node
    launch_fill<double>(reinterpret_cast<double *>(acc), 0.0, 2 * (int64_t)n, x->stream);
    HIPCHK(hipGetLastError());
    dft_install(x, comp, axis, p, ref_compat, interval, N, n, w, acc);
    if (bin) *bin = k;
    HIPCHK(hipStreamSynchronize(x->stream));
    return 0;
}

int phpfdtd_set_dft_noalloc(phpfdtd_ctx *x, int comp, char axis, int pos, int ref_compat, int interval,
                            int nsamples, double f_hz, int *bin) {
    return guarded(
        [&] { return set_dft_noalloc_impl(x, comp, axis, pos, ref_compat, interval, nsamples, f_hz, bin); });
}

static int read_dft_impl(phpfdtd_ctx *x, double *re, double *im, int *n_samples) {
    if (!x) return fail(PHPFDTD_EINVAL, "null ctx");
    if (!x->dft_on) return fail(PHPFDTD_ESTATE, "no DFT set (phpfdtd_set_dft)");
    HIPCHK(hipSetDevice(x->device));
    HIPCHK(hipStreamSynchronize(x->stream));
    const DftArgs &a = x->dft;
    if (n_samples) {
        const int64_t done = x->host_step - a.t0;  // steps t0 .. host_step-1 have run
        const int64_t ns = done <= 0 ? 0 : (done - 1) / a.interval + 1;
        *n_samples = (int)std::min<int64_t>(ns, a.nsamp);
    }
    if (!re && !im) return 0;
    std::vector<double2> h(a.nplane);
    HIPCHK(hipMemcpy(h.data(), a.acc, h.size() * sizeof(double2), hipMemcpyDeviceToHost));
    for (int64_t q = 0; q < a.nplane; ++q) {
        if (re) re[q] = h[q].x;
        if (im) im[q] = h[q].y;
    }
    return 0;
}

int phpfdtd_read_dft(phpfdtd_ctx *x, double *re, double *im, int *n_samples) {
    return guarded([&] { return read_dft_impl(x, re, im, n_samples); });
}

static int run_impl(phpfdtd_ctx *x, int nsteps) {
    if (!x || nsteps < 0) return fail(PHPFDTD_EINVAL, "bad arguments");
    if (x->batch) return fail(PHPFDTD_ESTATE, "context is a batch member; step it with phpfdtd_batch_run");
    if (x->host_step + nsteps > x->g.tsteps)
        return fail(PHPFDTD_EINVAL, "run past tsteps (%lld + %d > %d)", (long long)x->host_step, nsteps, x->g.tsteps);
    HIPCHK(hipSetDevice(x->device));
    if (x->dirty) {
        int rc = rebuild(x);
        if (rc) return rc;
    }
    int n = nsteps;
    if (x->use_graph) {
        while (n >= 16) {
            HIPCHK(hipGraphLaunch(x->g16, x->stream));
            n -= 16;
        }
        while (n-- > 0) HIPCHK(hipGraphLaunch(x->g1, x->stream));
    } else {
        while (n-- > 0)
            for (auto &k : x->plan) k.fn(x->stream);
        HIPCHK(hipGetLastError());
    }
    x->host_step += nsteps;
    return 0;
}

int phpfdtd_run(phpfdtd_ctx *x, int nsteps) {
    return guarded([&] { return run_impl(x, nsteps); });
}

static int sync_impl(phpfdtd_ctx *x, phpfdtd_diag *d) {
    if (!x) return fail(PHPFDTD_EINVAL, "null ctx");
    HIPCHK(hipSetDevice(x->device));
    HIPCHK(hipStreamSynchronize(x->stream));
    if (!d) return 0;
    unsigned long long init[4] = {0, 0, 0, ~0ull};
    HIPCHK(hipMemcpyAsync(x->d_diag, init, sizeof(init), hipMemcpyHostToDevice, x->stream));
    if (x->prec == PHPFDTD_FP64)
        launch_reduce_e<double>((double *)x->cur(0), (double *)x->cur(1), (double *)x->cur(2), x->nx, x->ny, x->nz, x->nxp,
                                x->d_diag, x->stream);
    else
        launch_reduce_e<float>((float *)x->cur(0), (float *)x->cur(1), (float *)x->cur(2), x->nx, x->ny, x->nz, x->nxp,
                               x->d_diag, x->stream);
    // one pass over each lattice material's r: max |r| and its first argmax
    for (auto &sp : x->specials)
        if (sp.pre.natom > 0)
            launch_diag_r(sp.d_r, x->pk, sp.d_cells, (int64_t)sp.cells.size(), sp.lay, x->ny, x->nz, x->nxp,
                          x->d_diag_part, x->d_diag, x->stream);
    unsigned long long res[4];
    HIPCHK(hipMemcpyAsync(res, x->d_diag, sizeof(res), hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    memset(d, 0, sizeof(*d));
    d->step = x->host_step;
    memcpy(&d->max_e, &res[0], 8);
    memcpy(&d->max_ex, &res[1], 8);
    memcpy(&d->max_r, &res[2], 8);
    d->r_i = d->r_j = d->r_k = d->r_atom = d->r_dir = -1;
    if (res[3] != ~0ull && d->max_r > 0) {
        const unsigned long long key = res[3];
        const int64_t dof = key % 96, cell = key / 96;
        d->r_k = (int)(cell % x->nz);
        d->r_j = (int)((cell / x->nz) % x->ny);
        d->r_i = (int)(cell / ((int64_t)x->nz * x->ny));
        d->r_atom = (int)(dof / 3);
        d->r_dir = (int)(dof % 3);
    }
    d->diverged = d->max_r > 1e-10;
    if (d->diverged)
        return fail(PHPFDTD_EDIVERGED, "ERROR: r field too large at step %lld: max r %e at i=%d, j=%d, k=%d atom %d dir %d",
                    (long long)x->host_step - 1, d->max_r, d->r_i, d->r_j, d->r_k, d->r_atom, d->r_dir);
    return 0;
}

int phpfdtd_sync(phpfdtd_ctx *x, phpfdtd_diag *d) {
    return guarded([&] { return sync_impl(x, d); });
}

static int read_slice_impl(phpfdtd_ctx *x, int comp, char axis, int pos, int ref_compat, float *out) {
    if (!x || !out || comp < 0 || comp > 9) return fail(PHPFDTD_EINVAL, "bad arguments");
    size_t n;
    int p;
    int rc = plane_of(x, axis, pos, ref_compat, &n, &p);
    if (rc) return rc;
    HIPCHK(hipSetDevice(x->device));
    if (axis == 'z' && p >= x->nz) {  // the reference reads its zero upper halo plane
        HIPCHK(hipStreamSynchronize(x->stream));
        memset(out, 0, n * sizeof(float));
        return 0;
    }
    float *buf = x->d_slice;
    const int emag = comp == PHPFDTD_EMAG;
    const int fc = emag ? 0 : comp;
    if (fc >= 3 && fc <= 5)
        derive_box(x, axis == 'x' ? p : 0, axis == 'x' ? p + 1 : x->nx, axis == 'y' ? p : 0, axis == 'y' ? p + 1 : x->ny,
                   axis == 'z' ? p : 0, axis == 'z' ? p + 1 : x->nz);
    if (x->prec == PHPFDTD_FP64)
        launch_slice<double>((double *)x->cur(fc), (double *)x->cur(0), (double *)x->cur(1), (double *)x->cur(2), emag, axis,
                             p, x->nx, x->ny, x->nz, x->nxp, (float *)buf, x->stream);
    else
        launch_slice<float>((float *)x->cur(fc), (float *)x->cur(0), (float *)x->cur(1), (float *)x->cur(2), emag, axis, p,
                            x->nx, x->ny, x->nz, x->nxp, (float *)buf, x->stream);
    HIPCHK(hipMemcpyAsync(out, buf, n * sizeof(float), hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    return 0;
}

int phpfdtd_read_slice(phpfdtd_ctx *x, int comp, char axis, int pos, int ref_compat, float *out) {
    return guarded([&] { return read_slice_impl(x, comp, axis, pos, ref_compat, out); });
}

static int read_monitor_impl(phpfdtd_ctx *x, double *rows, int max_rows, int *n_rows) {
    if (!x || !n_rows || max_rows < 0 || (max_rows > 0 && !rows)) return fail(PHPFDTD_EINVAL, "bad arguments");
    HIPCHK(hipSetDevice(x->device));
    *n_rows = 0;
    if (!x->mon_on) return 0;
    HIPCHK(hipStreamSynchronize(x->stream));
    const int64_t avail = std::min(x->host_step, x->ring_rows) - x->mon_read;
    const int64_t n = std::min<int64_t>(avail, max_rows);
    if (n <= 0) return 0;
    HIPCHK(hipMemcpy(rows, x->d_ring + x->mon_read * 12, n * 12 * sizeof(double), hipMemcpyDeviceToHost));
    x->mon_read += n;
    *n_rows = (int)n;
    return 0;
}

int phpfdtd_read_monitor(phpfdtd_ctx *x, double *rows, int max_rows, int *n_rows) {
    return guarded([&] { return read_monitor_impl(x, rows, max_rows, n_rows); });
}

static int read_point_impl(phpfdtd_ctx *x, int px, int py, int pz, double *f9) {
    if (!x || !f9 || px < 0 || px >= x->nx || py < 0 || py >= x->ny || pz < 0 || pz > x->nz)
        return fail(PHPFDTD_EINVAL, "bad arguments");
    HIPCHK(hipSetDevice(x->device));
    if (pz < x->nz) derive_box(x, px, px + 1, py, py + 1, pz, pz + 1);
    HIPCHK(hipStreamSynchronize(x->stream));
    for (int q = 0; q < 9; ++q) f9[q] = 0.0;
    if (pz == x->nz) return 0;
    const size_t c = x->fidx(px, py, pz);
    for (int q = 0; q < 9; ++q) {
        if (x->prec == PHPFDTD_FP64) {
            HIPCHK(hipMemcpy(&f9[q], (double *)x->cur(q) + c, 8, hipMemcpyDeviceToHost));
        } else {
            float v;
            HIPCHK(hipMemcpy(&v, (float *)x->cur(q) + c, 4, hipMemcpyDeviceToHost));
            f9[q] = v;
        }
    }
    return 0;
}

int phpfdtd_read_point(phpfdtd_ctx *x, int px, int py, int pz, double *f9) {
    return guarded([&] { return read_point_impl(x, px, py, pz, f9); });
}

static int read_ions_impl(phpfdtd_ctx *x, int px, int py, int pz, double *r, double *v) {
    if (!x || px < 0 || px >= x->nx || py < 0 || py >= x->ny || pz < 0 || pz >= x->nz)
        return fail(PHPFDTD_EINVAL, "bad arguments");
    HIPCHK(hipSetDevice(x->device));
    const size_t c = x->fidx(px, py, pz);
    const int m = x->medium[c];
    const int sp = x->special_of[m];
    if (sp < 0 || x->pre[m].natom == 0) return 0;
    int64_t slot = 0;
    if (find_slot(x, sp, c, &slot)) return fail(PHPFDTD_EINVAL, "cell has no slot");
    const Special &s = x->specials[sp];
    const size_t n = s.cells.size();
    const int n3 = 3 * x->pre[m].natom;
    HIPCHK(hipStreamSynchronize(x->stream));
    (void)n;
    for (int q = 0; q < n3; ++q) {
        for (int k = 0; k < 2; ++k) {
            double *out = k == 0 ? r : v;
            if (!out) continue;
            const size_t at = (size_t)(k == 0 ? s.lay.r_index(slot, q) : s.lay.v_index(slot, q));
            const char *dev = (const char *)(k == 0 ? s.d_r : s.d_v);
            if (x->pk) {
                float w;
                uint8_t b;
                bool in_r = k == 0;  // v's code may sit in r's array
                const int64_t ci = k == 0 ? s.lay.c_index(slot, q) : s.lay.cv_index(slot, q, &in_r);  // nibble index
                HIPCHK(hipMemcpy(&w, dev + at * 4, 4, hipMemcpyDeviceToHost));
                HIPCHK(hipMemcpy(&b, (const char *)(in_r ? s.d_r : s.d_v) + (ci >> 1), 1, hipMemcpyDeviceToHost));
                out[q] = pk_unpack(w, pk_code_at(&b, ci & 1));
            } else {
                HIPCHK(hipMemcpy(out + q, dev + at * 8, 8, hipMemcpyDeviceToHost));
            }
        }
    }
    return x->pre[m].natom;
}

int phpfdtd_read_ions(phpfdtd_ctx *x, int px, int py, int pz, double *r, double *v) {
    return guarded([&] { return read_ions_impl(x, px, py, pz, r, v); });
}

static int read_field_impl(phpfdtd_ctx *x, int comp, double *out) {
    if (!x || !out || comp < 0 || comp > 8) return fail(PHPFDTD_EINVAL, "bad arguments");
    HIPCHK(hipSetDevice(x->device));
    if (comp >= 3 && comp <= 5) derive_box(x, 0, x->nx, 0, x->ny, 0, x->nz);
    HIPCHK(hipStreamSynchronize(x->stream));
    std::vector<unsigned char> h(x->ncellp * x->rs);
    HIPCHK(hipMemcpy(h.data(), x->cur(comp), h.size(), hipMemcpyDeviceToHost));
    for (int i = 0; i < x->nx; ++i)
        for (int j = 0; j < x->ny; ++j)
            for (int k = 0; k < x->nz; ++k) {
                const size_t c = x->fidx(i, j, k);
                out[((size_t)i * x->ny + j) * x->nz + k] =
                    x->prec == PHPFDTD_FP64 ? ((double *)h.data())[c] : (double)((float *)h.data())[c];
            }
    return 0;
}

int phpfdtd_read_field(phpfdtd_ctx *x, int comp, double *out) {
    return guarded([&] { return read_field_impl(x, comp, out); });
}

static int write_field_impl(phpfdtd_ctx *x, int comp, const double *in) {
    if (!x || !in || comp < 0 || comp > 8) return fail(PHPFDTD_EINVAL, "bad arguments");
    HIPCHK(hipSetDevice(x->device));
    HIPCHK(hipStreamSynchronize(x->stream));
    std::vector<unsigned char> h(x->ncellp * x->rs, 0);
    for (int i = 0; i < x->nx; ++i)
        for (int j = 0; j < x->ny; ++j)
            for (int k = 0; k < x->nz; ++k) {
                const size_t c = x->fidx(i, j, k);
                const double v = in[((size_t)i * x->ny + j) * x->nz + k];
                if (x->prec == PHPFDTD_FP64) ((double *)h.data())[c] = v;
                else ((float *)h.data())[c] = (float)v;
            }
    // E and H go to both sets: cells the update never changes (H at i = nx-1
    // or j = ny-1, fdtd.c:3123-3124) then hold the same value in either
    HIPCHK(hipMemcpy(x->F[0][comp], h.data(), h.size(), hipMemcpyHostToDevice));
    if (x->F[1][comp] != x->F[0][comp])
        HIPCHK(hipMemcpy(x->F[1][comp], h.data(), h.size(), hipMemcpyHostToDevice));
    return 0;
}

int phpfdtd_write_field(phpfdtd_ctx *x, int comp, const double *in) {
    return guarded([&] { return write_field_impl(x, comp, in); });
}

static int count_cells_impl(phpfdtd_ctx *x, int mat, int64_t *n) {
    if (!x || !n || mat < 0 || mat >= x->nmat) return fail(PHPFDTD_EINVAL, "bad arguments");
    const int sp = x->special_of[mat];
    *n = sp >= 0 ? (int64_t)x->specials[sp].cells.size() : 0;
    return 0;
}

int phpfdtd_count_cells(phpfdtd_ctx *x, int mat, int64_t *n) {
    return guarded([&] { return count_cells_impl(x, mat, n); });
}

// order of a material's slots in file order (i slowest, k fastest)
static std::vector<int64_t> file_order(const phpfdtd_ctx *x, const Special &s) {
    std::vector<int64_t> idx(s.cells.size());
    std::vector<uint64_t> key(s.cells.size());
    for (size_t q = 0; q < s.cells.size(); ++q) {
        const uint64_t c = s.cells[q];
        const uint64_t i = c % x->nxp, j = (c / x->nxp) % x->ny, k = c / ((uint64_t)x->nxp * x->ny);
        key[q] = (i * x->ny + j) * x->nz + k;
        idx[q] = (int64_t)q;
    }
    std::sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) { return key[a] < key[b]; });
    return idx;
}

// ADE state of the FP32 form: the device holds P and J = P - P^{n-1} (the
// increment form of the recurrence); the ABI exchanges P and P^{n-1}.  A
// write of P keeps P^{n-1} (J moves with P), a write of P^{n-1} sets J.
static int ade_state_f32(phpfdtd_ctx *x, Special &s, int kind, double *buf, int write) {
    const size_t n = s.cells.size();
    std::vector<float> P(3 * n), J(3 * n);
    HIPCHK(hipStreamSynchronize(x->stream));
    HIPCHK(hipMemcpy(P.data(), s.d_P, P.size() * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(J.data(), s.d_Po, J.size() * 4, hipMemcpyDeviceToHost));
    const std::vector<int64_t> ord = file_order(x, s);
    for (size_t q = 0; q < n; ++q)
        for (int d = 0; d < 3; ++d) {
            const size_t i = (size_t)d * n + ord[q];
            double &u = buf[q * 3 + d];
            if (!write) {
                u = kind == 2 ? (double)P[i] : (double)P[i] - (double)J[i];
            } else if (kind == 2) {
                const double pold = (double)P[i] - (double)J[i];
                P[i] = (float)u;
                J[i] = (float)((double)P[i] - pold);
            } else {
                J[i] = (float)((double)P[i] - u);
            }
        }
    if (write) {
        HIPCHK(hipMemcpy(s.d_P, P.data(), P.size() * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(s.d_Po, J.data(), J.size() * 4, hipMemcpyHostToDevice));
    }
    return 0;
}

static int state_io(phpfdtd_ctx *x, int mat, int kind, double *buf, int write) {
    if (!x || !buf || mat < 0 || mat >= x->nmat || kind < 0 || kind > 3) return fail(PHPFDTD_EINVAL, "bad arguments");
    HIPCHK(hipSetDevice(x->device));
    const int sp = x->special_of[mat];
    if (sp < 0) return 0;
    Special &s = x->specials[sp];
    if (kind >= 2 && s.pre.ade && x->prec == PHPFDTD_FP32) return ade_state_f32(x, s, kind, buf, write);
    const size_t n = s.cells.size();
    int width;
    void *dev;
    bool pk = false;  // compensated lattice state: FP32 words and 4-bit compensation codes in one array
    size_t es;
    if (kind < 2) {
        if (!s.pre.natom) return fail(PHPFDTD_EINVAL, "material %d has no lattice", mat);
        width = 3 * s.pre.natom;
        dev = kind == 0 ? s.d_r : s.d_v;
        pk = x->pk;
        es = x->pk ? 4 : 8;
    } else {
        if (!s.pre.ade) return fail(PHPFDTD_EINVAL, "material %d has no ADE state", mat);
        width = 3;
        dev = kind == 2 ? s.d_P : s.d_Po;
        es = x->rs;
    }
    HIPCHK(hipStreamSynchronize(x->stream));
    const bool tiled = kind < 2;  // lattice r, v in 64-slot tiles; ADE P, Po as [3][nslot]
    const size_t nelem = tiled ? (n + 63) / 64 * (size_t)(kind == 0 ? s.lay.tile_r() : s.lay.tile_v()) : n * width;
    std::vector<unsigned char> h(nelem * es);
    const std::vector<int64_t> ord = file_order(x, s);
    if (!write || tiled) HIPCHK(hipMemcpy(h.data(), dev, h.size(), hipMemcpyDeviceToHost));
    // v's last codes in r's array (IonLayout::shared): r's array comes along
    const bool vshare = pk && kind == 1 && s.lay.shared();
    std::vector<unsigned char> hr(vshare ? (n + 63) / 64 * (size_t)s.lay.tile_r() * es : 0);
    if (vshare) HIPCHK(hipMemcpy(hr.data(), s.d_r, hr.size(), hipMemcpyDeviceToHost));
    for (size_t q = 0; q < n; ++q) {
        const int64_t slot = ord[q];
        for (int d = 0; d < width; ++d) {
            const size_t hi = tiled ? (size_t)(kind == 0 ? s.lay.r_index(slot, d) : s.lay.v_index(slot, d))
                                    : (size_t)d * n + slot;
            double &u = buf[q * width + d];
            if (pk) {  // compensated FP32: written values are rounded to 28 significant bits
                bool in_r = kind == 0;
                const int64_t ci = kind == 0 ? s.lay.c_index(slot, d) : s.lay.cv_index(slot, d, &in_r);  // nibble index
                unsigned char *cb = kind == 1 && in_r ? hr.data() : h.data();
                float w;
                int b;
                if (write) {
                    pk_pack(u, &w, &b);
                    ((float *)h.data())[hi] = w;
                    pk_code_set(cb, ci, b);
                } else {
                    u = pk_unpack(((float *)h.data())[hi], pk_code_at(cb, ci));
                }
            } else if (es == 8) {
                if (write) ((double *)h.data())[hi] = u;
                else u = ((double *)h.data())[hi];
            } else {
                if (write) ((float *)h.data())[hi] = (float)u;
                else u = ((float *)h.data())[hi];
            }
        }
    }
    if (write) HIPCHK(hipMemcpy(dev, h.data(), h.size(), hipMemcpyHostToDevice));
    if (write && vshare) HIPCHK(hipMemcpy(s.d_r, hr.data(), hr.size(), hipMemcpyHostToDevice));
    return 0;
}

int phpfdtd_read_state(phpfdtd_ctx *x, int mat, int kind, double *out) {
    return guarded([&] { return state_io(x, mat, kind, out, 0); });
}

static int write_state_impl(phpfdtd_ctx *x, int mat, int kind, const double *in) {
    return state_io(x, mat, kind, const_cast<double *>(in), 1);
}

int phpfdtd_write_state(phpfdtd_ctx *x, int mat, int kind, const double *in) {
    return guarded([&] { return write_state_impl(x, mat, kind, in); });
}

static int profile_impl(phpfdtd_ctx *x, int nsteps, phpfdtd_ktime *out, int max_out, int *n_out) {
    if (!x || nsteps <= 0 || !n_out) return fail(PHPFDTD_EINVAL, "bad arguments");
    if (x->host_step + nsteps > x->g.tsteps) return fail(PHPFDTD_EINVAL, "profile past tsteps");
    if (x->batch) return fail(PHPFDTD_ESTATE, "context is a batch member");
    HIPCHK(hipSetDevice(x->device));
    if (x->dirty) {
        int rc = rebuild(x);
        if (rc) return rc;
    }
    const size_t nk = x->plan.size();
    std::vector<hipEvent_t> ev((size_t)nsteps * nk * 2);
    for (auto &e : ev) HIPCHK(hipEventCreate(&e));
    for (int t = 0; t < nsteps; ++t)
        for (size_t k = 0; k < nk; ++k) {
            HIPCHK(hipEventRecord(ev[((size_t)t * nk + k) * 2], x->stream));
            x->plan[k].fn(x->stream);
            HIPCHK(hipEventRecord(ev[((size_t)t * nk + k) * 2 + 1], x->stream));
        }
    HIPCHK(hipStreamSynchronize(x->stream));
    x->host_step += nsteps;
    *n_out = 0;
    for (size_t k = 0; k < nk && (int)k < max_out; ++k) {
        double tot = 0;
        for (int t = 0; t < nsteps; ++t) {
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, ev[((size_t)t * nk + k) * 2], ev[((size_t)t * nk + k) * 2 + 1]));
            tot += ms;
        }
        phpfdtd_ktime &o = out[k];
        memset(&o, 0, sizeof(o));
        snprintf(o.name, sizeof(o.name), "%s", x->plan[k].name.c_str());
        o.avg_ms = tot / nsteps;
        o.bytes = x->plan[k].bytes;
        o.flops = x->plan[k].flops;
        o.launches = 1;
        *n_out = (int)k + 1;
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    return 0;
}

int phpfdtd_profile(phpfdtd_ctx *x, int nsteps, phpfdtd_ktime *out, int max_out, int *n_out) {
    return guarded([&] { return profile_impl(x, nsteps, out, max_out, n_out); });
}

static int time_steps_impl(phpfdtd_ctx *x, int nsteps, double *ms) {
    if (!x || !ms || nsteps <= 0) return fail(PHPFDTD_EINVAL, "bad arguments");
    if (x->batch) return fail(PHPFDTD_ESTATE, "context is a batch member");
    HIPCHK(hipSetDevice(x->device));
    if (x->dirty) {
        int rc = rebuild(x);
        if (rc) return rc;
    }
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    HIPCHK(hipStreamSynchronize(x->stream));
    HIPCHK(hipEventRecord(a, x->stream));
    int rc = phpfdtd_run(x, nsteps);
    if (rc) return rc;
    HIPCHK(hipEventRecord(b, x->stream));
    HIPCHK(hipEventSynchronize(b));
    float t = 0;
    HIPCHK(hipEventElapsedTime(&t, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *ms = t / nsteps;
    return 0;
}

int phpfdtd_time_steps(phpfdtd_ctx *x, int nsteps, double *ms) {
    return guarded([&] { return time_steps_impl(x, nsteps, ms); });
}

static int struct_sizes_impl(int *sizes, int n) {
    const int s[5] = {(int)sizeof(phpfdtd_grid), (int)sizeof(phpfdtd_material), (int)sizeof(phpfdtd_source),
                      (int)sizeof(phpfdtd_diag), (int)sizeof(phpfdtd_ktime)};
    for (int q = 0; q < n && q < 5; ++q) sizes[q] = s[q];
    return 5;
}

int phpfdtd_struct_sizes(int *sizes, int n) {
    return guarded([&] { return struct_sizes_impl(sizes, n); });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batches: one launch per kernel kind over the tiles of all member contexts.
// The per-launch tables (argument blocks, tile -> member map, first tile of
// each member) live in one device buffer, rebuilt when membership or a
// member's configuration changes.  The upload is stream-ordered on the batch
// stream, so it lands after every kernel already queued that reads the old
// tables; two pinned staging buffers let the host build the next tables while
// the previous upload is still pending.
struct phpfdtd_batch {
    int prec = 0;  // field form of the members (PHPFDTD_FP64 / PHPFDTD_FP32)
    int form = 0;  // the members' precision value (PHPFDTD_FP64 / FP32 / FP32_RV64)
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev = nullptr;
    std::vector<phpfdtd_ctx *> members;
    bool dirty = true;
    bool fused = true;  // Yee plan of the members: one-pass (8-row work-groups) or split
    // device tables
    unsigned char *d_tab = nullptr;
    size_t d_cap = 0;
    unsigned char *h_stage[2] = {nullptr, nullptr};
    size_t h_cap[2] = {0, 0};
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    int stage_next = 0;
    // launches of one step (pointers into d_tab)
    BatchLaunch<YeeArgs<double>> yee64{};
    BatchLaunch<YeeArgs<float>> yee32{};
    struct SlotGroup {
        int kind = 0;
        BatchLaunch<SlotArgs<double>> b64{};
        BatchLaunch<SlotArgs<float>> b32{};
    };
    std::vector<SlotGroup> groups;
    const StepEndArgs *d_se = nullptr;
    int max_nmat = 1;
    // completion events of queued phpfdtd_batch_run calls, oldest first
    std::vector<hipEvent_t> inflight, spare;

    ~phpfdtd_batch() {
        if (!stream) return;
        (void)hipSetDevice(device);
        (void)hipStreamSynchronize(stream);
        for (int q = 0; q < 2; ++q) {
            if (h_stage[q]) (void)hipHostFree(h_stage[q]);
            if (stage_ev[q]) (void)hipEventDestroy(stage_ev[q]);
        }
        if (d_tab) (void)hipFree(d_tab);
        if (ev) (void)hipEventDestroy(ev);
        for (hipEvent_t e : inflight) (void)hipEventDestroy(e);
        for (hipEvent_t e : spare) (void)hipEventDestroy(e);
        (void)hipStreamDestroy(stream);
    }
};

// Mark a context's plan stale, and its batch's tables with it.
static void touch(phpfdtd_ctx *x) {
    x->dirty = true;
    if (x->batch) x->batch->dirty = true;
}

namespace {
// Host image of the batch tables: byte blob with 256-byte aligned sections.
struct Blob {
    std::vector<unsigned char> b;
    size_t put(const void *p, size_t n) {
        const size_t off = (b.size() + 255) & ~(size_t)255;
        b.resize(off + n);
        if (n) memcpy(b.data() + off, p, n);
        return off;
    }
};

template <typename A>
struct LaunchImage {
    std::vector<A> args;
    std::vector<uint32_t> map;
    std::vector<int> start;
    size_t o_args = 0, o_map = 0, o_start = 0;
    int total = 0, total_pad = 0;
    void add(const A &a, int64_t tiles) {
        start.push_back((int)map.size());
        const uint32_t rid = (uint32_t)args.size();
        args.push_back(a);
        map.insert(map.end(), (size_t)tiles, rid);
    }
    void place(Blob &bl, int pad_to) {
        total = (int)map.size();
        total_pad = (total + pad_to - 1) / pad_to * pad_to;
        o_args = bl.put(args.data(), args.size() * sizeof(A));
        o_map = bl.put(map.data(), map.size() * sizeof(uint32_t));
        o_start = bl.put(start.data(), start.size() * sizeof(int));
    }
    BatchLaunch<A> bind(unsigned char *base) const {
        BatchLaunch<A> l;
        l.args = reinterpret_cast<const A *>(base + o_args);
        l.map = reinterpret_cast<const uint32_t *>(base + o_map);
        l.start = reinterpret_cast<const int *>(base + o_start);
        l.total = total;
        l.total_pad = total_pad;
        return l;
    }
};
}  // namespace

template <typename real>
static int batch_tables_t(phpfdtd_batch *b) {
    LaunchImage<YeeArgs<real>> yee;
    std::vector<int> kinds;
    std::vector<LaunchImage<SlotArgs<real>>> slot;
    std::vector<StepEndArgs> se;
    b->max_nmat = 1;
    for (phpfdtd_ctx *x : b->members) {
        build_plan(x);
        const YeeArgs<real> &a = args_of<real>(x);
        yee.add(a, a.nblocks);
        b->max_nmat = std::max(b->max_nmat, x->nmat);
        for (auto &sp : x->specials) {
            const SlotArgs<real> &sa = slot_args_of<real>(sp);
            const int kind = slot_kind(sa);
            size_t g = 0;
            while (g < kinds.size() && kinds[g] != kind) ++g;
            if (g == kinds.size()) {
                kinds.push_back(kind);
                slot.emplace_back();
            }
            slot[g].add(sa, slot_blocks(sa));
        }
        se.push_back(x->se);
    }
    Blob bl;
    yee.place(bl, 8);
    for (auto &l : slot) l.place(bl, 1);
    const size_t o_se = bl.put(se.data(), se.size() * sizeof(StepEndArgs));
    const size_t n = bl.b.size();
    // device buffer (grown only after the stream has drained the old one)
    if (n > b->d_cap) {
        HIPCHK(hipStreamSynchronize(b->stream));
        if (b->d_tab) HIPCHK(hipFree(b->d_tab));
        b->d_tab = nullptr;
        const size_t cap = std::max(n * 2, (size_t)1 << 20);
        hipError_t e = hipMalloc((void **)&b->d_tab, cap);
        if (e != hipSuccess) {
            b->d_cap = 0;
            return fail(PHPFDTD_ENOMEM_HBM, "batch tables: hipMalloc(%zu): %s", cap, hipGetErrorString(e));
        }
        b->d_cap = cap;
    }
    // pinned staging buffer whose previous upload has completed
    const int q = b->stage_next;
    b->stage_next ^= 1;
    HIPCHK(hipEventSynchronize(b->stage_ev[q]));
    if (n > b->h_cap[q]) {
        if (b->h_stage[q]) HIPCHK(hipHostFree(b->h_stage[q]));
        b->h_stage[q] = nullptr;
        const size_t cap = std::max(n * 2, (size_t)1 << 20);
        HIPCHK(hipHostMalloc((void **)&b->h_stage[q], cap, hipHostMallocDefault));
        b->h_cap[q] = cap;
    }
    memcpy(b->h_stage[q], bl.b.data(), n);
    HIPCHK(hipMemcpyAsync(b->d_tab, b->h_stage[q], n, hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipEventRecord(b->stage_ev[q], b->stream));
    if constexpr (sizeof(real) == 8) b->yee64 = yee.bind(b->d_tab);
    else b->yee32 = yee.bind(b->d_tab);
    b->groups.clear();
    for (size_t g = 0; g < kinds.size(); ++g) {
        phpfdtd_batch::SlotGroup sg;
        sg.kind = kinds[g];
        if constexpr (sizeof(real) == 8) sg.b64 = slot[g].bind(b->d_tab);
        else sg.b32 = slot[g].bind(b->d_tab);
        b->groups.push_back(sg);
    }
    b->d_se = reinterpret_cast<const StepEndArgs *>(b->d_tab + o_se);
    b->dirty = false;
    return 0;
}

template <typename real>
static void batch_step_t(phpfdtd_batch *b) {
    const BatchLaunch<YeeArgs<real>> &y = [&]() -> const BatchLaunch<YeeArgs<real>> & {
        if constexpr (sizeof(real) == 8) return b->yee64;
        else return b->yee32;
    }();
    if (b->fused) {
        launch_yee_batch<real>(y, b->max_nmat, b->stream);
    } else {
        launch_h_batch<real>(y, b->stream);
        launch_de_batch<real>(y, b->max_nmat, b->stream);
    }
    for (auto &g : b->groups) {
        if constexpr (sizeof(real) == 8) launch_slot_batch<real>(g.b64, g.kind, b->stream);
        else launch_slot_batch<real>(g.b32, g.kind, b->stream);
    }
    for (phpfdtd_ctx *x : b->members)
        if (x->dft_on) launch_dft<real>(x->dft, b->stream);
    launch_step_end_batch<real>(b->d_se, (int)b->members.size(), b->stream);
}

extern "C" {

static int batch_create_impl(int precision, phpfdtd_batch **out) {
    if (!out || (precision != PHPFDTD_FP64 && precision != PHPFDTD_FP32 && precision != PHPFDTD_FP32_RV64))
        return fail(PHPFDTD_EINVAL, "bad arguments");
    *out = nullptr;
    std::unique_ptr<phpfdtd_batch> b(new phpfdtd_batch());
    b->prec = precision == PHPFDTD_FP64 ? PHPFDTD_FP64 : PHPFDTD_FP32;
    b->form = precision;
    HIPCHK(hipGetDevice(&b->device));
    const char *by = getenv("PHPFDTD_BATCH_YEE");
    b->fused = !(by && strcmp(by, "split") == 0);
    HIPCHK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&b->ev, hipEventDisableTiming));
    for (int q = 0; q < 2; ++q) {
        HIPCHK(hipEventCreateWithFlags(&b->stage_ev[q], hipEventDisableTiming));
        HIPCHK(hipEventRecord(b->stage_ev[q], b->stream));
    }
    *out = b.release();
    return 0;
}

int phpfdtd_batch_create(int precision, phpfdtd_batch **out) {
    return guarded([&] { return batch_create_impl(precision, out); });
}

static int batch_add_impl(phpfdtd_batch *b, phpfdtd_ctx *x) {
    if (!b || !x) return fail(PHPFDTD_EINVAL, "null argument");
    if (x->batch) return fail(PHPFDTD_ESTATE, "context is already in a batch");
    if (x->g.precision != b->form)
        return fail(PHPFDTD_EINVAL, "context precision %d differs from the batch's %d", x->g.precision, b->form);
    if (x->device != b->device) return fail(PHPFDTD_EINVAL, "context is on device %d, batch on %d", x->device, b->device);
    // The batch's Yee plan (same results either way): the one-pass kernel
    // with 8-row work-groups and 16-plane chunks, or the split pair with
    // 4-plane FP32 chunks.  The member's own plan returns when it leaves the
    // batch (its next rebuild).
    x->fused = b->fused;
    x->yee_rows = kBatchRows;
    x->yee_kchunk = x->prec == PHPFDTD_FP32 && !b->fused ? 4 : 16;
    HIPCHK(hipSetDevice(x->device));
    if (!x->ev) HIPCHK(hipEventCreateWithFlags(&x->ev, hipEventDisableTiming));
    b->members.push_back(x);
    x->batch = b;
    b->dirty = true;
    return 0;
}

int phpfdtd_batch_add(phpfdtd_batch *b, phpfdtd_ctx *x) {
    return guarded([&] { return batch_add_impl(b, x); });
}

static int batch_remove_impl(phpfdtd_batch *b, phpfdtd_ctx *x) {
    if (!b || !x) return fail(PHPFDTD_EINVAL, "null argument");
    auto it = std::find(b->members.begin(), b->members.end(), x);
    if (it == b->members.end()) return fail(PHPFDTD_EINVAL, "context is not a member of this batch");
    b->members.erase(it);
    x->batch = nullptr;
    x->dirty = true;  // its graphs were captured before joining or are stale
    b->dirty = true;
    return 0;
}

int phpfdtd_batch_remove(phpfdtd_batch *b, phpfdtd_ctx *x) {
    return guarded([&] { return batch_remove_impl(b, x); });
}

static int batch_run_impl(phpfdtd_batch *b, int nsteps) {
    if (!b || nsteps < 0) return fail(PHPFDTD_EINVAL, "bad arguments");
    for (phpfdtd_ctx *x : b->members)
        if (x->host_step + nsteps > x->g.tsteps)
            return fail(PHPFDTD_EINVAL, "member runs past tsteps (%lld + %d > %d)", (long long)x->host_step, nsteps,
                        x->g.tsteps);
    if (b->members.empty() || nsteps == 0) return 0;
    HIPCHK(hipSetDevice(b->device));
    if (b->dirty) {
        int rc = b->prec == PHPFDTD_FP64 ? batch_tables_t<double>(b) : batch_tables_t<float>(b);
        if (rc) return rc;
    }
    // members' own queued work (uploads, state writes) comes first
    for (phpfdtd_ctx *x : b->members) {
        HIPCHK(hipEventRecord(x->ev, x->stream));
        HIPCHK(hipStreamWaitEvent(b->stream, x->ev, 0));
    }
    for (int t = 0; t < nsteps; ++t) {
        if (b->prec == PHPFDTD_FP64) batch_step_t<double>(b);
        else batch_step_t<float>(b);
    }
    HIPCHK(hipGetLastError());
    // later work on a member's stream (reads, syncs) waits for these steps
    HIPCHK(hipEventRecord(b->ev, b->stream));
    hipEvent_t done;
    if (!b->spare.empty()) {
        done = b->spare.back();
        b->spare.pop_back();
    } else {
        HIPCHK(hipEventCreateWithFlags(&done, hipEventDisableTiming));
    }
    HIPCHK(hipEventRecord(done, b->stream));
    b->inflight.push_back(done);
    for (phpfdtd_ctx *x : b->members) {
        HIPCHK(hipStreamWaitEvent(x->stream, b->ev, 0));
        x->host_step += nsteps;
    }
    return 0;
}

int phpfdtd_batch_run(phpfdtd_batch *b, int nsteps) {
    return guarded([&] { return batch_run_impl(b, nsteps); });
}

static int batch_sync_impl(phpfdtd_batch *b) {
    if (!b) return fail(PHPFDTD_EINVAL, "null batch");
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    for (hipEvent_t e : b->inflight) b->spare.push_back(e);
    b->inflight.clear();
    return 0;
}

int phpfdtd_batch_sync(phpfdtd_batch *b) {
    return guarded([&] { return batch_sync_impl(b); });
}

static int batch_wait_impl(phpfdtd_batch *b, int keep) {
    if (!b || keep < 0) return fail(PHPFDTD_EINVAL, "bad arguments");
    HIPCHK(hipSetDevice(b->device));
    while ((int)b->inflight.size() > keep) {
        HIPCHK(hipEventSynchronize(b->inflight.front()));
        b->spare.push_back(b->inflight.front());
        b->inflight.erase(b->inflight.begin());
    }
    return 0;
}

int phpfdtd_batch_wait(phpfdtd_batch *b, int keep) {
    return guarded([&] { return batch_wait_impl(b, keep); });
}

int phpfdtd_batch_size(const phpfdtd_batch *b) { return b ? (int)b->members.size() : 0; }

void phpfdtd_batch_destroy(phpfdtd_batch *b) {
    if (!b) return;
    for (phpfdtd_ctx *x : b->members) {
        x->batch = nullptr;
        x->dirty = true;
    }
    delete b;
}

}  // extern "C"
