// phpfdtd_internal.h - device data layout and kernel parameter blocks of
// libphpfdtd_hip.  Not part of the public ABI (include/phpfdtd.h).
//
// Device layout (HBM), chosen for gfx950 rather than copied from fdtd.c:
//   * fields Ex..Hz: one array per component, x fastest, row pitch nxp
//     (nx rounded up to 32 elements), cell (i,j,k) at (k*ny + j)*nxp + i.  A
//     lane of the Yee kernels reads a 16-byte vector of consecutive x cells,
//     a 64-lane wavefront a 1 KiB row segment.
//   * material ids: uint8 per cell, same layout.
//   * CPML psi: compact slabs, not full volumes:
//       x slabs [nz][ny][2L], y slabs [nz][2L][nxp], z slabs [2L][ny][nxp].
//   * ADE and lattice state: per-material slot arrays, slot s of material m
//     is the s-th cell of m in x-fastest order; ADE P[3][slot]; lattice r, v
//     in 64-slot tiles r[tile][dof / w][64][w] (IonLayout).
// FP64 host precompute (CPML b/c/kappa, eps^-1, ADE coefficients, rotations,
// source values) is cast to the working type only when uploaded.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "../../include/phpfdtd.h"

namespace phpfdtd {

constexpr int kMaxMat = 64;     // material ids are stored as uint8
constexpr int kMaxSrc = 8;      // sources handled inside the D/E kernel
constexpr int kMaxAtom = 32;    // lattice atoms per cell
constexpr int kPackedKind = 64; // batch kind offset of compensated-state lattice kernels
constexpr int kBlockY = 4;      // wavefronts per work-group of the split-plan Yee kernels (64 x 4 threads)
constexpr int kBatchRows = 8;   // y rows (wavefronts) per work-group of the batched one-pass Yee kernel
constexpr int kDiagWaves = 8192;  // wavefront partials of the one-pass lattice diagnostics
constexpr double kEpsInf = 20.0;  // EPS_INF of the reference ADE (fdtd.c:202)

// x cells per lane in the Yee kernels: one 16-byte vector.
template <typename real> struct YeeVec { static constexpr int value = 16 / sizeof(real); };

enum MatFlags : int {
    MF_LOSS = 1,     // eps_im != 0 and no ADE: D -= Lk . E^n (fdtd.c:3349-3353)
    MF_SPECIAL = 2,  // E is produced by a slot kernel (ADE and/or lattice)
    MF_DSTATE = 4,   // FP32 form: the cell keeps D in HBM (slot cells, and plain
                     // materials whose M has no inverse); other cells carry E alone
};

// Update form of the Yee kernels.  FP64 (the plain form) keeps the
// reference's D-form for every cell: D += curl H (+ psi, loss, sources), E =
// M . D (fdtd.c:3238-3462).  The FP32 precision form runs the same update
// in E-form for plain cells: with dD the increment above, E^{n+1} = E^n +
// M . dD, which equals M . (D^n + dD) since M is constant in time; their D
// is then not stored (it is M^-1 E, derived when read).  That removes D's
// read and write (6 words of 18) from every plain cell.  Cells with
// MF_DSTATE keep D += dD, and the slot kernels read it.
template <typename real>
struct EForm {
    static constexpr bool value = sizeof(real) == 4;
};

// Per-material coefficients read by the fused D/E kernel.
template <typename real>
struct MatCoef {
    real M[9];   // plain cells: E_lab = M . D_lab, M = R^T (lf (1-lt) Re(eps^-1)/eps0) R
    real Lk[9];  // D loss: D -= Lk . E^n_lab, Lk = dt f 2 3.14159265 eps0 eps_im
    int flags;
    int pad[3];
};

// One CPML axis: coefficient tables over the axis index (length n).
template <typename real>
struct Axis {
    const real *b;    // exp(-(sigma/(kappa eps0) + alpha/eps0) dt)
    const real *kh;   // H curl coefficient dt / (mu0 delta kappa)
    const real *kd;   // D curl coefficient dt / (delta kappa)
    const real *ch;   // H psi coefficient dt c / (mu0 delta)
    const real *cd;   // D psi coefficient dt c / delta
    const int *slot;  // psi slab index 0..2L-1, or -1 outside the CPML
};

// E and H are double-buffered: step s (the device step counter at the start
// of the step) reads set s & 1 and writes set (s + 1) & 1, so a work-group
// that recomputes a neighbour's H from old values never meets one that
// another work-group has already replaced; after s steps set s & 1 is
// current.  The H psi slabs follow H.  D and the D psi are read and written
// by their own cell only and stay single.
// CPML coefficients of one y row or z plane, packed so that one scalar load
// brings all of them (same meaning as the Axis tables).
template <typename real>
struct alignas(16) AxRow {
    real kh, kd, b, ch, cd;
    int p;  // psi slab index, -1 outside the CPML
};

template <typename real>
struct YeeArgs {
    real *E[2][3], *H[2][3];        // [set][x, y, z]
    real *Dx, *Dy, *Dz;
    real *pH[2][6];                 // [set][xy, xz (x slabs), yx, yz (y slabs), zx, zy (z slabs)]
    real *pDxy, *pDxz, *pDyx, *pDyz, *pDzx, *pDzy;
    const uint8_t *medium;
    const MatCoef<real> *mats;
    Axis<real> ax;                 // x: per-cell tables, read as lane vectors
    const AxRow<real> *ay, *az;    // y rows, z planes
    int nx, ny, nz, nxp, L;
    int nmat;  // entries of mats
    int kbase, kend;  // z planes [kbase, kend) of this launch
    int kchunk, nbx, nby, nbz, nblocks, nblocks_pad;
    int rows;  // y rows (wavefronts) per work-group
    // sources: cell and component (i, j, k, 0|1|2) in src[4s..4s+3], value*dt
    // per step in src_val[step*nsrc + s]
    int nsrc;
    const int *src;
    const real *src_val;
    const int64_t *step;
    uint64_t dmask;  // bit m: material m has MF_DSTATE
};

// MFMA tiling of the lattice contraction -DMM . r for N atoms: k-tiles of 4
// dofs and m-tiles of 16 rows over the 3N dofs.
__host__ __device__ inline int mfma_nk(int na) { return (3 * na + 3) / 4; }
__host__ __device__ inline int mfma_nma(int na) { return (3 * na + 15) / 16; }

// Work-group b of a single-domain Yee launch -> its tile: groups of g
// consecutive tiles (g = the x segments of a band of rows; tiles are numbered
// x segment fastest) are dealt to the XCDs in turn (work-group b runs on XCD
// b mod 8).  The launch runs a multiple of 8 g work-groups, so every tile
// has one; the rest get -1.  Measured on the 1024x1024x640 FP32 step:
// 7.71-7.73 ms against 8.35-8.41 ms for one contiguous range of tiles per
// XCD, 8.21-8.28 ms for tile b, 8.08-8.12 ms for four bands per group.
__host__ __device__ inline int band_tile(int b, int nblocks, int g) {
    const int x = b & 7, i = b >> 3;
    const int t = ((i / g) * 8 + x) * g + i % g;
    return t < nblocks ? t : -1;
}

// natom of the lattice kernel instantiation (register arrays sized for it;
// the generic one runs to the material's own counts).
inline int mfma_inst(int natom) {
    switch (natom) {
        case 1: case 2: case 3: case 4: case 6: case 8: case 12: case 16: return natom;
        default: return kMaxAtom;
    }
}

// Lattice state layout of one material: tiles of 64 slots, the 3N dofs of a
// slot taken in groups of w, one 16-byte vector per lane and group, so every
// wavefront access is one 1 KiB row.  3N is padded to a multiple of w with
// zero dofs.
//   * FP64 state (plain form, and the precision form with FP64 r, v): w = 2,
//     r[tile][dof / 2][64][2] doubles.
//   * compensated FP32 state (precision form): w = 4; a tile is its ng groups
//     of 1 KiB (the 64 lanes' four FP32 words), then ceil(ng / 8) blocks of
//     1 KiB of 4-bit compensation codes: word 4 lane + d of block b holds the
//     codes of that lane's groups 8b + 2d (bits 0-15) and 8b + 2d + 1 (bits
//     16-31), four bits per value, so the codes also move as one 16-byte
//     vector per lane.  Words and codes of an array stream from one address
//     range.  When the last block of r holds 1-4 groups (ng % 8 in 1..4: 16
//     atoms, 4 atoms ...) its words 2 and 3 of each lane are free and hold
//     the codes of v's last groups instead, so v's tile has one block fewer
//     (shared(); r and v are advanced together, so the block is loaded and
//     stored once per step).
constexpr int kPkWords = 256;  // 4-byte words per 1 KiB row

struct IonLayout {
    int n3;
    int w;  // dofs per lane vector: 2 (FP64) or 4 (compensated FP32)
    __host__ __device__ int n3p() const { return (n3 + w - 1) / w * w; }
    __host__ __device__ int ng() const { return n3p() / w; }    // dof groups (1 KiB rows) per tile
    __host__ __device__ int nb() const { return (ng() + 7) / 8; }  // compensation rows of r (w = 4)
    // v's last groups keep their codes in the free half of r's last row
    __host__ __device__ bool shared() const { return w == 4 && ng() % 8 != 0 && ng() % 8 <= 4; }
    __host__ __device__ int nbv() const { return nb() - (shared() ? 1 : 0); }  // compensation rows of v
    // elements per tile: doubles (FP64), or 4-byte words (compensated)
    __host__ __device__ int64_t tile_r() const { return w == 4 ? (int64_t)kPkWords * (ng() + nb()) : 64LL * n3p(); }
    __host__ __device__ int64_t tile_v() const { return w == 4 ? (int64_t)kPkWords * (ng() + nbv()) : 64LL * n3p(); }
    // element of dof q of slot s: the double, or the FP32 word
    __host__ __device__ int64_t r_index(int64_t s, int q) const {
        return (s >> 6) * tile_r() + (int64_t)(q / w) * 64 * w + (s & 63) * w + q % w;
    }
    __host__ __device__ int64_t v_index(int64_t s, int q) const {
        return (s >> 6) * tile_v() + (int64_t)(q / w) * 64 * w + (s & 63) * w + q % w;
    }
    // 4-bit code of dof q of slot s (compensated state), as a nibble index into
    // r's array (two per byte, the low one first): group j = q / 4 sits in
    // dword (j % 8) / 2 of the lane's vector in row j / 8, bits 16 (j % 2) + 4 (q % 4)
    __host__ __device__ int64_t c_index(int64_t s, int q) const {
        const int j = q / 4;
        const int64_t dw = (s >> 6) * tile_r() + (int64_t)kPkWords * (ng() + j / 8) + (s & 63) * 4 + (j % 8) / 2;
        return dw * 8 + 4 * (j % 2) + q % 4;
    }
    // the same for v: a nibble index into v's array, or, for the groups of a
    // shared last row (*in_r set), into r's array at dwords 2 + (j % 8) / 2
    __host__ __device__ int64_t cv_index(int64_t s, int q, bool *in_r) const {
        const int j = q / 4;
        *in_r = shared() && j / 8 == nb() - 1;
        const int64_t dw = *in_r ? (s >> 6) * tile_r() + (int64_t)kPkWords * (ng() + j / 8) + (s & 63) * 4 + 2 + (j % 8) / 2
                                 : (s >> 6) * tile_v() + (int64_t)kPkWords * (ng() + j / 8) + (s & 63) * 4 + (j % 8) / 2;
        return dw * 8 + 4 * (j % 2) + q % 4;
    }
    // slot and dof of element idx; false for padding and compensation words
    __host__ __device__ bool r_decode(int64_t idx, int64_t nslot, int64_t *s, int *q) const {
        const int64_t t = idx / tile_r(), rem = idx % tile_r();
        const int gsz = 64 * w;
        if (rem >= (int64_t)gsz * ng()) return false;
        const int u = (int)(rem % gsz);
        *q = (int)(rem / gsz) * w + u % w;
        *s = t * 64 + u / w;
        return *q < n3 && *s < nslot;
    }
};

inline IonLayout ion_layout(int natom, bool packed = false) {
    IonLayout l;
    l.n3 = 3 * natom;
    l.w = packed ? 4 : 2;
    return l;
}

// Compensated FP32 value (hi, lo): x = hi + lo * 2^(e - 154), e the biased
// exponent of hi clamped to >= 1, so lo counts units of 2^-4 of hi's ulp and
// the 4-bit code lo in [-7, 7] covers the rounding error of hi = fl32(x)
// (|x - hi| <= ulp / 2; a code of +-8 is clamped to +-7, so a stored value
// stays strictly inside half an ulp of hi and re-packs to itself).  The
// stored value carries 28 significant bits: pack and unpack are exact in FP64
// apart from that final rounding of x - hi to a multiple of 2^-4 ulp.
__host__ __device__ inline int pk_exp(float hi) {
    uint32_t b;
    __builtin_memcpy(&b, &hi, 4);
    const int e = (int)((b >> 23) & 0xffu);
    return e > 1 ? e : 1;
}
// 2^k as a double, for the exponent range pk_* uses (-157 .. 157)
__host__ __device__ inline double pk_pow2(int k) {
    const uint64_t b = (uint64_t)(k + 1023) << 52;
    double d;
    __builtin_memcpy(&d, &b, 8);
    return d;
}
__host__ __device__ inline double pk_unpack(float hi, int lo) {
    return fma((double)lo, pk_pow2(pk_exp(hi) - 154), (double)hi);  // exact
}
__host__ __device__ inline void pk_pack(double x, float *hi, int *lo) {
    const float h = (float)x;
    // (x - h) in units of 2^-4 ulp(h), |.| <= 8, rounded to nearest even by
    // adding 1.5 * 2^52: the integer lands in the low word of the sum
    const double q = (x - (double)h) * pk_pow2(154 - pk_exp(h)) + 6755399441055744.0;
    uint64_t b;
    __builtin_memcpy(&b, &q, 8);
    const int n = (int)(uint32_t)b;
    *hi = h;
    *lo = n < -7 ? -7 : n > 7 ? 7 : n;  // a non-finite x (diverged state) gives any code
}
// The 4-bit code at nibble index n of an array (two per byte, the low first),
// sign-extended; and its replacement.
__host__ __device__ inline int pk_code_at(const void *base, int64_t n) {
    const uint8_t b = static_cast<const uint8_t *>(base)[n >> 1];
    const int v = (n & 1) ? (b >> 4) : (b & 15);
    return (v ^ 8) - 8;
}
__host__ inline void pk_code_set(void *base, int64_t n, int c) {
    uint8_t &b = static_cast<uint8_t *>(base)[n >> 1];
    b = (n & 1) ? (uint8_t)((b & 0x0f) | ((c & 15) << 4)) : (uint8_t)((b & 0xf0) | (c & 15));
}
// The code at bit s of a dword of a compensation row, and its replacement.
__host__ __device__ inline int pk_code_of(uint32_t w, int s) { return ((int)(w << (28 - s))) >> 28; }
__host__ __device__ inline uint32_t pk_with_code(uint32_t w, int s, int c) {
    return (w & ~(0xfu << s)) | ((uint32_t)(c & 15) << s);
}

// Lattice / ADE slot kernels: one material per launch.
template <typename real>
struct SlotArgs {
    real *E[2][3];            // double-buffered E (YeeArgs): read set step & 1, write the other
    const real *Dx, *Dy, *Dz;
    const int64_t *step;
    const uint32_t *cells;  // slot -> field index
    int64_t nslot;
    // crystal-frame coefficients
    real R[9];    // lab -> crystal rotation
    real A[9];    // E_cr = A . (D_cr - P)  (lf, 1-lt, Re(eps^-1), 1/eps0 folded)
    // ADE (fdtd.c:3463-3479), P^{n+1} = Gpi (2P - Gm P^{n-1} + W E^n_cr).
    // FP64 form: that recurrence on (P, P^{n-1}).  FP32 form: the same
    // recurrence on (P, J = P - P^{n-1}): u = J - Ga J + W E, J' = u - Gb u,
    // P' = P + J' (Ga = gamma dt/2, Gb = I - Gpi, both small), which keeps
    // the marginal P mode exact in FP32.
    int ade;
    real Gpi[9], Gm[9], W[9];
    real Ga[9], Gb[9];
    real *P, *Po;  // [3][nslot]
    // lattice (fdtd.c:3505-3528)
    int natom;
    void *r, *v;             // lattice state, layout `lay`
    int pk;                  // 1: compensated FP32 state (lay.w == 4), 0: FP64
    IonLayout lay;
    const double *qezv;      // [N][3][3]
    const double *qezm;      // [N][3][3]
    const double *dmm;       // [3N][3N]
    const float *aop;        // precision form: -DMM as MFMA A operands in lane order
    const float *qop;        // precision form: [QeZV | QeZM] in FP32 (9N each)
    double dt, half_dt2;
};

// Per-step epilogue: monitor row + step counter.  When the next step's H
// update overlaps the lattice kernel, the row is written in two parts: D and
// H (part 1) before the overlap, t, E, R, V and the step advance (part 2)
// after it.
struct StepEndArgs {
    int64_t *step;
    double *ring;      // [tsteps][12]
    int64_t ring_rows;
    int mon_on;
    int64_t mon_c;     // field index sampled (z+1 in reference mode), -1 = zero halo
    const void *F[2][9];  // field arrays (real) of both E/H sets; the D entries are the same in both
    const void *mon_r, *mon_v;    // lattice state of the monitor cell's material
    int64_t mon_ridx, mon_vidx;   // elements of (slot, atom*3+dir) in r and v, -1 = no R/V
    int64_t mon_rc, mon_vc;       // nibble indices of their 4-bit codes (compensated state), else -1
    const void *mon_vcb;          // array holding v's code: v, or r for a shared last row
    double dt;
    int part;                     // 0 whole row, 1 D and H, 2 t, E, R, V + step advance
    int mon_dfe;                  // E-form plain monitor cell: D = mon_dm . E (D is not stored)
    double mon_dm[9];
};

// D = Dm . E for the E-form plain cells of a box of cells (their D is not
// stored), written to the D arrays so that readers find it there.  The E
// set read is (step + ahead) & 1: ahead 1 inside a step's plan before the
// counter advances, 0 between steps.
struct DeriveArgs {
    const void *E[2][3];
    void *D[3];
    const uint8_t *medium;
    const double *dm;             // [nmat][9] D-from-E matrix (M^-1), FP64
    const uint8_t *derive;        // [nmat] 1 = E-form plain material
    int nxp, ny;
    int i0, i1, j0, j1, k0, k1;   // box [i0, i1) x [j0, j1) x [k0, k1)
    int ahead;
    const int64_t *step;
};

// One launch over the work of several domains (a batch): tile g of the grid
// belongs to domain map[g] and is tile g - start[map[g]] of that domain.
template <typename A>
struct BatchLaunch {
    const A *args;          // [ndomain] argument blocks
    const uint32_t *map;    // [total]
    const int *start;       // [ndomain]
    int total, total_pad;   // tiles, and grid size (multiple of 8 for the Yee kernels)
};

// Wavefront-uniform reads (coefficient tables, source values, batch tables)
// go through the scalar cache: the constant address space is read-only for
// the kernel's lifetime, so these become s_load and wait on their own counter
// instead of queueing behind the field streams in the vector-memory counter.
template <typename T>
__device__ __forceinline__ T sld(const T *p, int64_t idx) {
    typedef const __attribute__((address_space(4))) T *cptr;
    return ((cptr)p)[idx];
}

// Argument block of domain rid of a batched launch (wavefront-uniform, so it
// is loaded through the scalar cache into scalar registers).
template <typename A>
__device__ __forceinline__ A batch_args(const BatchLaunch<A> &b, int rid) {
    static_assert(sizeof(A) % 4 == 0, "argument blocks are copied as dwords");
    typedef const __attribute__((address_space(4))) uint32_t *cptr;
    const cptr src = (cptr)(b.args + rid);
    A out;
    uint32_t *dst = reinterpret_cast<uint32_t *>(&out);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(A) / 4); ++i) dst[i] = src[i];
    return out;
}

// Streamed single-bin DFT of one plane (phpfdtd_set_dft).
struct DftArgs {
    const void *F[2], *Ex[2], *Ey[2], *Ez[2];  // component sampled (real), and E for |E|, per E/H set
    int emag, ref_float;
    char axis;
    int pos;                       // plane index actually read (z+1 applied)
    int nx, ny, nz, nxp;
    int64_t ncell;                 // cells of a field array (plane past the top reads 0)
    int64_t nplane;                // elements of the plane
    int64_t t0, interval, nsamp;
    const double2 *w;              // [nsamp] exp(-2 pi i k n / N) / N
    double2 *acc;                  // [nplane]
    const int64_t *step;
};

}  // namespace phpfdtd
