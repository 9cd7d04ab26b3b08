// kernels_yee.hip - the Yee update of one time step for gfx950: H, D and the
// E of plain cells (fdtd.c:3123-3462).
//
// Every lane owns a 16-byte vector of consecutive x cells (4 in FP32, 2 in
// FP64), so a wavefront covers a 1 KiB x segment of one y row and every field
// stream moves in 16-byte lane accesses.  A wavefront marches a short chunk
// of z planes, carrying the z neighbours in registers; x neighbours are the
// adjacent element or a cross-lane shuffle.  Hardware places work-group b on
// XCD (b mod 8); band_tile() renumbers the groups of one domain so that the
// x segments of a band of rows run on one XCD (sharing its L2 for the x halo
// cells) and consecutive bands go to the eight XCDs in turn, which keeps all
// of them streaming the same region of each plane.  Batches use xcd_tile(),
// one contiguous range of tiles per XCD.
//
// E and H are double-buffered (YeeArgs): step s reads set s & 1 and writes
// set (s + 1) & 1.  D and the D psi are updated in place.
//
// Two plans give the same bits:
//   * fused (yee_update, the default): one pass.  At each plane a wavefront
//     computes the new H of its cells, then their new D and the E of plain
//     cells.  The R wavefronts of a work-group (R = 8 by default) hold R
//     consecutive rows and pass E and the new H of their rows to the
//     neighbouring wavefront through LDS.  D needs the new H of neighbours outside the work-group -
//     row j0-1, plane k0-1 at the start of the chunk, and cell x0-1 for lane
//     0 - which it recomputes from the old E, H and psi with the same code
//     (h_cells) their owner runs; the old values cannot have been replaced by
//     then, since new values go to the other set.  E, H and D each move once
//     per cell and step: 18 words plus the material id.
//   * split (h_update, then de_update): the new H is written, then read back
//     by the D/E kernel: 21 words per cell.  Batched launches (many small
//     domains in one grid) run the one-pass tile too (yee_update_batch),
//     or this plan on request.
// The arithmetic is written with explicit fused multiply-adds and the file is
// compiled without FP contraction, so the owner's H and every recomputation
// of it round identically.
//
// Reference semantics (fdtd.c):
//   H: i < nx-1, j < ny-1, all k; E above the top plane is 0 (3123-3231).
//   D: 1 <= i <= nx-2, 1 <= j <= ny-2, all k; H below k=0 is 0 (3238-3356).
//   CPML psi is lagged: the old psi enters the field, then psi is advanced;
//   psi is stored pre-multiplied by dt.  H and psi of cells outside the loops
//   are never changed.
//   D loss uses E^n in the lab frame with unrotated eps_im (3349-3353).
//   Source: D[comp] += value*dt after the curl and loss (3363-3396), on any row.
//   E: every cell (3404-3462); cells of ADE/lattice materials are left to the
//   slot kernels (kernels_slot.hip).
#include "phpfdtd_internal.h"

namespace phpfdtd {

__device__ __forceinline__ int xcd_tile(int b, int nblocks, int nblocks_pad) {
    const int per = nblocks_pad >> 3;
    const int t = (b & 7) * per + (b >> 3);
    return t < nblocks ? t : -1;
}


__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

template <typename real>
struct Vec;
template <>
struct Vec<float> {
    static constexpr int n = 4;
    typedef float type __attribute__((ext_vector_type(4)));
    typedef unsigned int ids;  // 4 material ids
};
template <>
struct Vec<double> {
    static constexpr int n = 2;
    typedef double type __attribute__((ext_vector_type(2)));
    typedef unsigned short ids;  // 2 material ids
};

template <typename V, typename T>
__device__ __forceinline__ V ld(const T *p) {
    return *reinterpret_cast<const V *>(p);
}
template <typename V, typename T>
__device__ __forceinline__ void st(V v, T *p) {
    *reinterpret_cast<V *>(p) = v;
}
// streams touched once per step
template <typename V, typename T>
__device__ __forceinline__ V ldnt(const T *p) {
    return __builtin_nontemporal_load(reinterpret_cast<const V *>(p));
}
template <typename V, typename T>
__device__ __forceinline__ void stnt(V v, T *p) {
    __builtin_nontemporal_store(v, reinterpret_cast<V *>(p));
}

template <typename V>
__device__ __forceinline__ V vfma(V a, V b, V c) {
    return __builtin_elementwise_fma(a, b, c);
}

template <typename real>
__device__ __forceinline__ real shfl_down1(real v) {
    return __shfl_down(v, 1, 64);
}
template <typename real>
__device__ __forceinline__ real shfl_up1(real v) {
    return __shfl_up(v, 1, 64);
}

// The vector of the i+1 neighbours: elements 1.. of v, then `next`.
template <typename V, typename real>
__device__ __forceinline__ V next_x(V v, real next) {
    constexpr int n = sizeof(V) / sizeof(real);
    V w;
#pragma unroll
    for (int e = 0; e < n - 1; ++e) w[e] = v[e + 1];
    w[n - 1] = next;
    return w;
}
// The vector of the i-1 neighbours: `prev`, then elements ..n-2 of v.
template <typename V, typename real>
__device__ __forceinline__ V prev_x(V v, real prev) {
    constexpr int n = sizeof(V) / sizeof(real);
    V w;
    w[0] = prev;
#pragma unroll
    for (int e = 1; e < n; ++e) w[e] = v[e - 1];
    return w;
}

enum : int { HX = 1, HY = 2, HZ = 4 };

// CPML coefficients of one y row or z plane, one scalar load.
template <typename real>
using AxC = AxRow<real>;
template <typename real>
__device__ __forceinline__ AxC<real> axc(const AxRow<real> *t, int i) {
    static_assert(sizeof(AxRow<real>) % 4 == 0, "rows are read as dwords");
    typedef const __attribute__((address_space(4))) uint32_t *cptr;
    const cptr src = (cptr)(t + i);
    AxC<real> c;
    uint32_t *dst = reinterpret_cast<uint32_t *>(&c);
#pragma unroll
    for (int w = 0; w < (int)(sizeof(AxRow<real>) / 4); ++w) dst[w] = src[w];
    return c;
}

// The VN x cells of one lane vector, first cell x0.  The x coefficient
// tables (padded to nxp) are read where they are used, so they hold no
// registers across the z march.
template <typename real>
struct LaneX {
    int sl0;                   // x psi slab index of the first cell, -1 outside the CPML
    int xr;                    // in-row offset (lanes past the padded row read cell 0)
    bool vx;                   // FP32: the cells are consecutive slots of one aligned slab vector
    bool hpsi, dpsi;           // a cell with an x psi slot inside the H / D loop
    bool hon[Vec<real>::n];    // inside the H loop: i < nx-1
    bool don[Vec<real>::n];    // inside the D loop: 1 <= i <= nx-2
};
template <typename real>
__device__ __forceinline__ LaneX<real> lane_x(const YeeArgs<real> &a, int x0) {
    typedef typename Vec<real>::type V;
    constexpr int VN = Vec<real>::n;
    LaneX<real> l;
    l.xr = x0 < a.nxp ? x0 : 0;
    const int sl0 = a.ax.slot[l.xr];
    l.sl0 = sl0;
    bool vx = true;
    l.hpsi = l.dpsi = false;
#pragma unroll
    for (int e = 0; e < VN; ++e) {
        const int sl = a.ax.slot[l.xr + e];
        const int i = x0 + e;
        vx = vx && sl == sl0 + e;
        l.hon[e] = i < a.nx - 1;
        l.don[e] = i >= 1 && i <= a.nx - 2;
        l.hpsi |= sl >= 0 && l.hon[e];
        l.dpsi |= sl >= 0 && l.don[e];
    }
    // the vector psi path: FP32 only (2.9% off the FP32 step, neutral in FP64
    // where it costs registers)
    l.vx = vx && sizeof(real) == 4 && (2 * a.L) % VN == 0 && sl0 >= 0 && sl0 % VN == 0;
    return l;
}

// E and H sets of this step: o read (set step & 1), n written.
template <typename real>
struct Sets {
    const real *Eo[3], *Ho[3], *pHo[6];
    real *En[3], *Hn[3], *pHn[6];
};
template <typename real>
__device__ __forceinline__ Sets<real> sets(const YeeArgs<real> &a, int64_t step) {
    const bool odd = step & 1;
    Sets<real> s;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        s.Eo[q] = odd ? a.E[1][q] : a.E[0][q];
        s.En[q] = odd ? a.E[0][q] : a.E[1][q];
        s.Ho[q] = odd ? a.H[1][q] : a.H[0][q];
        s.Hn[q] = odd ? a.H[0][q] : a.H[1][q];
    }
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        s.pHo[q] = odd ? a.pH[1][q] : a.pH[0][q];
        s.pHn[q] = odd ? a.pH[0][q] : a.pH[1][q];
    }
    return s;
}

// New H (MASK: components) of lane vector lx in row j, plane k, from the old
// E, H and psi (fdtd.c:3123-3231).  E inputs: e? at (j, k); exk, eyk at
// (j, k+1); exj, ezj at (j+1, k); eyi, ezi at the i+1 neighbours.  With
// WRITE (the cells' owner, lane holding cells: `store`) the advanced psi goes
// to set n; a recomputation by a neighbour writes nothing.  Cells outside the
// H loop keep H and psi.  The psi terms enter in x, y, z order.
template <typename real, int MASK, bool WRITE>
__device__ __forceinline__ void h_cells(const YeeArgs<real> &a, const Sets<real> &s, const LaneX<real> &lx,
                                        const AxC<real> &yc, const AxC<real> &zc, int j, int k, bool store,
                                        typename Vec<real>::type ex, typename Vec<real>::type ey,
                                        typename Vec<real>::type ez, typename Vec<real>::type exk,
                                        typename Vec<real>::type eyk, typename Vec<real>::type exj,
                                        typename Vec<real>::type ezj, typename Vec<real>::type eyi,
                                        typename Vec<real>::type ezi, typename Vec<real>::type &hx,
                                        typename Vec<real>::type &hy, typename Vec<real>::type &hz) {
    typedef typename Vec<real>::type V;
    constexpr int VN = Vec<real>::n;
    const size_t nxp = a.nxp;
    const int L2 = 2 * a.L;
    const V dyEz = ezj - ez, dzEy = eyk - ey, dxEz = ezi - ez;
    const V dzEx = exk - ex, dxEy = eyi - ey, dyEx = exj - ex;
    V nhx = hx, nhy = hy, nhz = hz;
    const V khx = (MASK & (HY | HZ)) ? ld<V>(a.ax.kh + lx.xr) : V(0);
    if (MASK & HX) nhx = vfma(V(-yc.kh), dyEz, vfma(V(zc.kh), dzEy, hx));
    if (MASK & HY) nhy = vfma(V(-zc.kh), dzEx, vfma(khx, dxEz, hy));
    if (MASK & HZ) nhz = vfma(-khx, dxEy, vfma(V(yc.kh), dyEx, hz));
    if (MASK & (HY | HZ)) {  // x slabs: psi_Hxy (Hy), psi_Hxz (Hz)
        if (lx.vx) {
            const size_t q = ((size_t)k * a.ny + j) * L2 + lx.sl0;
            const V b = ld<V>(a.ax.b + lx.xr), c = ld<V>(a.ax.ch + lx.xr);
            if (MASK & HY) {
                const V o = ld<V>(s.pHo[0] + q);
                nhy += o;
                V n = vfma(b, o, c * dxEz);
#pragma unroll
                for (int e = 0; e < VN; ++e)
                    if (!lx.hon[e]) n[e] = o[e];
                if (WRITE && store) st(n, s.pHn[0] + q);
            }
            if (MASK & HZ) {
                const V o = ld<V>(s.pHo[1] + q);
                nhz += o;
                V n = vfma(b, o, -(c * dxEy));
#pragma unroll
                for (int e = 0; e < VN; ++e)
                    if (!lx.hon[e]) n[e] = o[e];
                if (WRITE && store) st(n, s.pHn[1] + q);
            }
        } else if (lx.hpsi) {
#pragma unroll
            for (int e = 0; e < VN; ++e) {
                const int i = lx.xr + e;
                const int px = a.ax.slot[i];
                if (px < 0 || !lx.hon[e]) continue;
                const size_t q = ((size_t)k * a.ny + j) * L2 + px;
                const real b = a.ax.b[i], c = a.ax.ch[i];
                if (MASK & HY) {
                    const real o = s.pHo[0][q];
                    nhy[e] += o;
                    if (WRITE) s.pHn[0][q] = fma(b, o, c * dxEz[e]);
                }
                if (MASK & HZ) {
                    const real o = s.pHo[1][q];
                    nhz[e] += o;
                    if (WRITE) s.pHn[1][q] = fma(b, o, -(c * dxEy[e]));
                }
            }
        }
    }
    if ((MASK & (HX | HZ)) && yc.p >= 0) {  // y slabs: psi_Hyx (Hx), psi_Hyz (Hz)
        const size_t q = ((size_t)k * L2 + yc.p) * nxp + lx.xr;
        if (MASK & HX) {
            const V o = ld<V>(s.pHo[2] + q);
            nhx += o;
            V n = vfma(V(yc.b), o, -(V(yc.ch) * dyEz));
#pragma unroll
            for (int e = 0; e < VN; ++e)
                if (!lx.hon[e]) n[e] = o[e];
            if (WRITE && store) st(n, s.pHn[2] + q);
        }
        if (MASK & HZ) {
            const V o = ld<V>(s.pHo[3] + q);
            nhz += o;
            V n = vfma(V(yc.b), o, V(yc.ch) * dyEx);
#pragma unroll
            for (int e = 0; e < VN; ++e)
                if (!lx.hon[e]) n[e] = o[e];
            if (WRITE && store) st(n, s.pHn[3] + q);
        }
    }
    if ((MASK & (HX | HY)) && zc.p >= 0) {  // z slabs: psi_Hzx (Hx), psi_Hzy (Hy)
        const size_t q = ((size_t)zc.p * a.ny + j) * nxp + lx.xr;
        if (MASK & HX) {
            const V o = ld<V>(s.pHo[4] + q);
            nhx += o;
            V n = vfma(V(zc.b), o, V(zc.ch) * dzEy);
#pragma unroll
            for (int e = 0; e < VN; ++e)
                if (!lx.hon[e]) n[e] = o[e];
            if (WRITE && store) st(n, s.pHn[4] + q);
        }
        if (MASK & HY) {
            const V o = ld<V>(s.pHo[5] + q);
            nhy += o;
            V n = vfma(V(zc.b), o, -(V(zc.ch) * dzEx));
#pragma unroll
            for (int e = 0; e < VN; ++e)
                if (!lx.hon[e]) n[e] = o[e];
            if (WRITE && store) st(n, s.pHn[5] + q);
        }
    }
#pragma unroll
    for (int e = 0; e < VN; ++e)
        if (!lx.hon[e]) {
            nhx[e] = hx[e];
            nhy[e] = hy[e];
            nhz[e] = hz[e];
        }
    hx = nhx;
    hy = nhy;
    hz = nhz;
}

// New Hy, Hz of the single cell (i, j, k), i < nx-1, from the old E, H and
// psi: what h_cells computes for that element (each vector operation there
// rounds per element exactly as the scalar one here), for lane 0's i-1
// neighbour.  ezi, eyi: Ez, Ey of cell i+1.
template <typename real>
__device__ __forceinline__ void h_cell_yz(const YeeArgs<real> &a, const Sets<real> &s, const AxC<real> &yc,
                                          const AxC<real> &zc, int i, int j, int k, size_t c, bool top,
                                          real ezi, real eyi, real &hy, real &hz) {
    const size_t nxp = a.nxp;
    const size_t plane = nxp * a.ny;
    const int L2 = 2 * a.L;
    const real ex = s.Eo[0][c], ey = s.Eo[1][c], ez = s.Eo[2][c];
    const real exk = top ? real(0) : s.Eo[0][c + plane];
    const real exj = s.Eo[0][c + nxp];
    const real kh = a.ax.kh[i];
    const real dxEz = ezi - ez, dzEx = exk - ex, dxEy = eyi - ey, dyEx = exj - ex;
    real nhy = fma(-zc.kh, dzEx, fma(kh, dxEz, s.Ho[1][c]));
    real nhz = fma(-kh, dxEy, fma(yc.kh, dyEx, s.Ho[2][c]));
    const int px = a.ax.slot[i];
    if (px >= 0) {
        const size_t q = ((size_t)k * a.ny + j) * L2 + px;
        nhy += s.pHo[0][q];
        nhz += s.pHo[1][q];
    }
    if (yc.p >= 0) nhz += s.pHo[3][((size_t)k * L2 + yc.p) * nxp + i];
    if (zc.p >= 0) nhy += s.pHo[5][((size_t)zc.p * a.ny + j) * nxp + i];
    hy = nhy;
    hz = nhz;
}

// New D of lane vector lx in row j (inside the D loop), plane k
// (fdtd.c:3238-3353): curl of the new H (h? at the cells, hxj, hzj at row
// j-1, hxk, hyk at plane k-1, hyi, hzi at the i-1 neighbours), psi advanced
// in place, then the eps_im loss with E^n (e?) for lossy cells.  Cells
// outside the D loop keep D.
template <typename real>
__device__ __forceinline__ void d_cells(const YeeArgs<real> &a, const LaneX<real> &lx, const AxC<real> &yc,
                                        const AxC<real> &zc, int j, int k, bool store, const MatCoef<real> *mats,
                                        const int *mid, const int *fl, bool anyloss, typename Vec<real>::type hx,
                                        typename Vec<real>::type hy, typename Vec<real>::type hz,
                                        typename Vec<real>::type hxj, typename Vec<real>::type hzj,
                                        typename Vec<real>::type hxk, typename Vec<real>::type hyk,
                                        typename Vec<real>::type hyi, typename Vec<real>::type hzi,
                                        typename Vec<real>::type ex, typename Vec<real>::type ey,
                                        typename Vec<real>::type ez, typename Vec<real>::type &dx,
                                        typename Vec<real>::type &dy, typename Vec<real>::type &dz) {
    typedef typename Vec<real>::type V;
    constexpr int VN = Vec<real>::n;
    const size_t nxp = a.nxp;
    const int L2 = 2 * a.L;
    const V dyHz = hz - hzj, dzHy = hy - hyk, dzHx = hx - hxk;
    const V dxHz = hz - hzi, dxHy = hy - hyi, dyHx = hx - hxj;
    const V kdx = ld<V>(a.ax.kd + lx.xr);
    V ndx = vfma(V(-zc.kd), dzHy, vfma(V(yc.kd), dyHz, dx));
    V ndy = vfma(-kdx, dxHz, vfma(V(zc.kd), dzHx, dy));
    V ndz = vfma(V(-yc.kd), dyHx, vfma(kdx, dxHy, dz));
    if (lx.vx) {  // x slabs: psi_Dxy (Dy), psi_Dxz (Dz)
        const size_t q = ((size_t)k * a.ny + j) * L2 + lx.sl0;
        const V b = ld<V>(a.ax.b + lx.xr), c = ld<V>(a.ax.cd + lx.xr);
        const V oy = ld<V>(a.pDxy + q), oz = ld<V>(a.pDxz + q);
        ndy += oy;
        ndz += oz;
        V ny = vfma(b, oy, -(c * dxHz)), nz = vfma(b, oz, c * dxHy);
#pragma unroll
        for (int e = 0; e < VN; ++e)
            if (!lx.don[e]) {
                ny[e] = oy[e];
                nz[e] = oz[e];
            }
        if (store) {
            st(ny, a.pDxy + q);
            st(nz, a.pDxz + q);
        }
    } else if (lx.dpsi) {
#pragma unroll
        for (int e = 0; e < VN; ++e) {
            const int i = lx.xr + e;
            const int px = a.ax.slot[i];
            if (px < 0 || !lx.don[e]) continue;
            const size_t q = ((size_t)k * a.ny + j) * L2 + px;
            const real b = a.ax.b[i], c = a.ax.cd[i];
            const real oy = a.pDxy[q], oz = a.pDxz[q];
            ndy[e] += oy;
            ndz[e] += oz;
            a.pDxy[q] = fma(b, oy, -(c * dxHz[e]));
            a.pDxz[q] = fma(b, oz, c * dxHy[e]);
        }
    }
    if (yc.p >= 0) {  // y slabs: psi_Dyx (Dx), psi_Dyz (Dz)
        const size_t q = ((size_t)k * L2 + yc.p) * nxp + lx.xr;
        const V ox = ld<V>(a.pDyx + q), oz = ld<V>(a.pDyz + q);
        ndx += ox;
        ndz += oz;
        V nx = vfma(V(yc.b), ox, V(yc.cd) * dyHz), nz = vfma(V(yc.b), oz, -(V(yc.cd) * dyHx));
#pragma unroll
        for (int e = 0; e < VN; ++e)
            if (!lx.don[e]) {
                nx[e] = ox[e];
                nz[e] = oz[e];
            }
        if (store) {
            st(nx, a.pDyx + q);
            st(nz, a.pDyz + q);
        }
    }
    if (zc.p >= 0) {  // z slabs: psi_Dzx (Dx), psi_Dzy (Dy)
        const size_t q = ((size_t)zc.p * a.ny + j) * nxp + lx.xr;
        const V ox = ld<V>(a.pDzx + q), oy = ld<V>(a.pDzy + q);
        ndx += ox;
        ndy += oy;
        V nx = vfma(V(zc.b), ox, -(V(zc.cd) * dzHy)), ny = vfma(V(zc.b), oy, V(zc.cd) * dzHx);
#pragma unroll
        for (int e = 0; e < VN; ++e)
            if (!lx.don[e]) {
                nx[e] = ox[e];
                ny[e] = oy[e];
            }
        if (store) {
            st(nx, a.pDzx + q);
            st(ny, a.pDzy + q);
        }
    }
    if (anyloss) {
#pragma unroll
        for (int e = 0; e < VN; ++e) {
            if (!(fl[e] & MF_LOSS)) continue;
            const real *Lk = mats[mid[e]].Lk;
            ndx[e] -= fma(Lk[0], ex[e], fma(Lk[1], ey[e], Lk[2] * ez[e]));
            ndy[e] -= fma(Lk[3], ex[e], fma(Lk[4], ey[e], Lk[5] * ez[e]));
            ndz[e] -= fma(Lk[6], ex[e], fma(Lk[7], ey[e], Lk[8] * ez[e]));
        }
    }
#pragma unroll
    for (int e = 0; e < VN; ++e)
        if (lx.don[e]) {
            dx[e] = ndx[e];
            dy[e] = ndy[e];
            dz[e] = ndz[e];
        }
}

// Soft sources of row j, plane k at step `st` (fdtd.c:3363-3396), added in
// source order.  Each element is taken as a scalar: an element index that
// depends on the source position would make the compiler index the vectors
// dynamically, through private memory.
template <typename real>
__device__ __forceinline__ void inject(const YeeArgs<real> &a, int x0, int j, int k, int64_t st,
                                       typename Vec<real>::type &dx, typename Vec<real>::type &dy,
                                       typename Vec<real>::type &dz) {
    constexpr int VN = Vec<real>::n;
    if (a.nsrc == 0) return;
#pragma unroll
    for (int q = 0; q < VN; ++q) {
        real x = dx[q], y = dy[q], z = dz[q];
        for (int s = 0; s < a.nsrc; ++s) {
            const int si = sld(a.src, 4 * s), sj = sld(a.src, 4 * s + 1), sk = sld(a.src, 4 * s + 2),
                      sc = sld(a.src, 4 * s + 3);
            if (sk == k && sj == j && si == x0 + q) {
                const real v = sld(a.src_val, st * a.nsrc + s);
                if (sc == 0) x += v;
                else if (sc == 1) y += v;
                else z += v;
            }
        }
        dx[q] = x;
        dy[q] = y;
        dz[q] = z;
    }
}

// D back in place, and E = M . D into set n for plain cells (fdtd.c:3404-3462).
template <typename real>
__device__ __forceinline__ void store_de(const YeeArgs<real> &a, const Sets<real> &s, size_t c,
                                         const MatCoef<real> *mats, const int *mid, const int *fl, bool allplain,
                                         typename Vec<real>::type dx, typename Vec<real>::type dy,
                                         typename Vec<real>::type dz) {
    typedef typename Vec<real>::type V;
    constexpr int VN = Vec<real>::n;
    stnt(dx, a.Dx + c);
    stnt(dy, a.Dy + c);
    stnt(dz, a.Dz + c);
    if (allplain) {
        V nex, ney, nez;
#pragma unroll
        for (int e = 0; e < VN; ++e) {
            const real *M = mats[mid[e]].M;
            nex[e] = fma(M[0], dx[e], fma(M[1], dy[e], M[2] * dz[e]));
            ney[e] = fma(M[3], dx[e], fma(M[4], dy[e], M[5] * dz[e]));
            nez[e] = fma(M[6], dx[e], fma(M[7], dy[e], M[8] * dz[e]));
        }
        stnt(nex, s.En[0] + c);
        stnt(ney, s.En[1] + c);
        stnt(nez, s.En[2] + c);
    } else {
#pragma unroll
        for (int e = 0; e < VN; ++e) {
            if (fl[e] & MF_SPECIAL) continue;
            const real *M = mats[mid[e]].M;
            s.En[0][c + e] = fma(M[0], dx[e], fma(M[1], dy[e], M[2] * dz[e]));
            s.En[1][c + e] = fma(M[3], dx[e], fma(M[4], dy[e], M[5] * dz[e]));
            s.En[2][c + e] = fma(M[6], dx[e], fma(M[7], dy[e], M[8] * dz[e]));
        }
    }
}

// E-form store (FP32, plain cells carry E alone): E^{n+1} = E^n + M . dD,
// the increment rounded on its own before it is added.  Cells with
// MF_DSTATE advance D (dx, dy, dz, loaded when the lane has one) by dD
// instead, and those that are not slot cells take E = M . D^{n+1}.
template <typename real>
__device__ __forceinline__ void store_ef(const YeeArgs<real> &a, const Sets<real> &s, size_t c,
                                         const MatCoef<real> *mats, const int *mid, const int *fl, bool anyd,
                                         typename Vec<real>::type ex, typename Vec<real>::type ey,
                                         typename Vec<real>::type ez, typename Vec<real>::type dx,
                                         typename Vec<real>::type dy, typename Vec<real>::type dz,
                                         typename Vec<real>::type gx, typename Vec<real>::type gy,
                                         typename Vec<real>::type gz) {
    typedef typename Vec<real>::type V;
    constexpr int VN = Vec<real>::n;
    if (!anyd) {
        V nex, ney, nez;
#pragma unroll
        for (int e = 0; e < VN; ++e) {
            const real *M = mats[mid[e]].M;
            nex[e] = ex[e] + fma(M[0], gx[e], fma(M[1], gy[e], M[2] * gz[e]));
            ney[e] = ey[e] + fma(M[3], gx[e], fma(M[4], gy[e], M[5] * gz[e]));
            nez[e] = ez[e] + fma(M[6], gx[e], fma(M[7], gy[e], M[8] * gz[e]));
        }
        stnt(nex, s.En[0] + c);
        stnt(ney, s.En[1] + c);
        stnt(nez, s.En[2] + c);
        return;
    }
    dx += gx;
    dy += gy;
    dz += gz;
    stnt(dx, a.Dx + c);
    stnt(dy, a.Dy + c);
    stnt(dz, a.Dz + c);
#pragma unroll
    for (int e = 0; e < VN; ++e) {
        if (fl[e] & MF_SPECIAL) continue;
        const real *M = mats[mid[e]].M;
        real nx, ny, nz;
        if (fl[e] & MF_DSTATE) {
            nx = fma(M[0], dx[e], fma(M[1], dy[e], M[2] * dz[e]));
            ny = fma(M[3], dx[e], fma(M[4], dy[e], M[5] * dz[e]));
            nz = fma(M[6], dx[e], fma(M[7], dy[e], M[8] * dz[e]));
        } else {
            nx = ex[e] + fma(M[0], gx[e], fma(M[1], gy[e], M[2] * gz[e]));
            ny = ey[e] + fma(M[3], gx[e], fma(M[4], gy[e], M[5] * gz[e]));
            nz = ez[e] + fma(M[6], gx[e], fma(M[7], gy[e], M[8] * gz[e]));
        }
        s.En[0][c + e] = nx;
        s.En[1][c + e] = ny;
        s.En[2][c + e] = nz;
    }
}

// Whether any of the lane's cells keeps D (E-form): bit id of the
// per-domain material mask, for each id byte of the word.
template <typename real>
__device__ __forceinline__ bool any_dstate(const YeeArgs<real> &a, typename Vec<real>::ids idw) {
    constexpr int VN = Vec<real>::n;
    bool any = false;
#pragma unroll
    for (int e = 0; e < VN; ++e) any |= ((a.dmask >> ((idw >> (8 * e)) & 0xff)) & 1) != 0;
    return any;
}

// Material ids of the lane's cells (one word) and their flags.
template <typename real>
__device__ __forceinline__ typename Vec<real>::ids cell_ids(const YeeArgs<real> &a, size_t c) {
    return *reinterpret_cast<const typename Vec<real>::ids *>(a.medium + c);
}
template <typename real>
__device__ __forceinline__ void cell_flags(typename Vec<real>::ids idw, const MatCoef<real> *mats, int *mid, int *fl,
                                           bool &anyloss, bool &allplain) {
    constexpr int VN = Vec<real>::n;
    anyloss = false;
    allplain = true;
#pragma unroll
    for (int e = 0; e < VN; ++e) {
        mid[e] = (idw >> (8 * e)) & 0xff;
        fl[e] = mats[mid[e]].flags;
        anyloss |= (fl[e] & MF_LOSS) != 0;
        allplain &= !(fl[e] & MF_SPECIAL);
    }
}
template <typename real>
__device__ __forceinline__ void cell_mats(const YeeArgs<real> &a, size_t c, const MatCoef<real> *mats, int *mid,
                                          int *fl, bool &anyloss, bool &allplain) {
    cell_flags<real>(cell_ids(a, c), mats, mid, fl, anyloss, allplain);
}

// Tile t of the grid: x segment bx (64 lane vectors), row j, z chunk bz.
struct TileOf {
    int bx, j, bz;
};
template <typename real>
__device__ __forceinline__ TileOf tile_of(const YeeArgs<real> &a, int t) {
    TileOf r;
    r.bx = t % a.nbx;
    const int rest = t / a.nbx;
    r.j = uniform((rest % a.nby) * a.rows + threadIdx.y);
    r.bz = rest / a.nby;
    return r;
}

// Fused plan: H, then D and E, plane by plane, for tile t of one domain.
// The ROWS wavefronts of a work-group hold ROWS consecutive rows and trade
// neighbour rows through LDS at one barrier per plane: a wavefront posts E of
// plane k+1 (row j+1's H input at the next plane, for the wavefront below)
// and its new H of plane k (row j-1's D input, for the wavefront above).
// Wavefront 0 recomputes the new H of row j0-1 and the last wavefront loads
// E of row j0+ROWS; every slot is double-buffered by plane parity.
template <typename real, int ROWS>
struct YeeRows {
    typename Vec<real>::type ex[2][ROWS][64], ez[2][ROWS][64];  // E of plane k (parity k & 1)
    typename Vec<real>::type hx[2][ROWS][64], hz[2][ROWS][64];  // new H of plane k
};
// The argument block in the constant address space (kernel arguments).
template <typename real>
using cYA = const __attribute__((address_space(4))) YeeArgs<real> *;
// `a` is the argument block and `base` its address in the constant address
// space.  Inside the plane loop the block is read again through `base`, made
// opaque to the compiler once per plane, so the fifty-odd pointers and sizes
// it holds are scalar loads next to their use rather than scalar registers
// held across the loop (which spilled: 172 spilled scalars, 490 lane reads
// in the loop; 50 and 132 this way, and 127 instead of 141 vector registers
// in FP64, 4 wavefronts per SIMD instead of 3).
template <typename real, int ROWS>
__device__ __forceinline__ void yee_tile(const YeeArgs<real> &a, cYA<real> base, const int t, const MatCoef<real> *mats) {
    typedef typename Vec<real>::type V;
    constexpr int VN = Vec<real>::n;
    __shared__ YeeRows<real, ROWS> rows;
    const TileOf tl = tile_of(a, t);
    const int j = tl.j;
    const int w = threadIdx.y;
    const int lane = threadIdx.x;
    const int x0 = (tl.bx * 64 + lane) * VN;
    if (uniform(tl.bx * 64 * VN) >= a.nx) return;  // whole work-group
    const bool live = j < a.ny;  // rows past the grid keep the barriers only
    const bool store = live && x0 < a.nx;  // the lane holds cells (lanes past the row end still load and shuffle)
    const int k0 = uniform(a.kbase + tl.bz * a.kchunk);
    const int k1 = min(k0 + a.kchunk, a.kend);
    const size_t nxp = a.nxp;
    const size_t plane = nxp * a.ny;
    const int64_t st = sld(a.step, 0);
    const Sets<real> s = sets(a, st);
    const LaneX<real> lx = lane_x(a, x0);
    const bool hrow = live && j < a.ny - 1;              // H loop row
    const bool drow = live && j >= 1 && j <= a.ny - 2;   // D loop row
    const bool first = w == 0, last = w == ROWS - 1;
    const int jj = live ? j : 0;
    const AxC<real> yc = axc(a.ay, jj);
    const AxC<real> yh = axc(a.ay, drow ? j - 1 : jj);
    const bool edge = lane == 63 && x0 + VN < a.nxp;  // loads the cell after its vector
    const bool xhalo = lane == 0 && x0 >= VN;         // recomputes the new H of cell x0-1

    size_t c = ((size_t)k0 * a.ny + jj) * nxp + lx.xr;
    V ex = V(0), ey = V(0), ez = V(0);
    V eyh = V(0);              // Ey of row j-1 at plane k (first wavefront)
    V hxk = V(0), hyk = V(0);  // new Hx, Hy at plane k-1 (zero below k = 0)
    if (live) {
        ex = ld<V>(s.Eo[0] + c), ey = ld<V>(s.Eo[1] + c), ez = ld<V>(s.Eo[2] + c);
        if (drow) {
            if (first) eyh = ld<V>(s.Eo[1] + c - nxp);
            if (k0 > 0) {  // new Hx, Hy of (j, k0-1)
                const size_t m = c - plane;
                const V exm = ld<V>(s.Eo[0] + m), eym = ld<V>(s.Eo[1] + m), ezm = ld<V>(s.Eo[2] + m);
                const V ezmj = ld<V>(s.Eo[2] + m + nxp);
                real n = shfl_down1(ezm[0]);
                if (edge) n = s.Eo[2][m + VN];
                V hz0 = V(0);
                hxk = ld<V>(s.Ho[0] + m);
                hyk = ld<V>(s.Ho[1] + m);
                h_cells<real, HX | HY, false>(a, s, lx, yc, axc(a.az, k0 - 1), j, k0 - 1, false, exm, eym, ezm, ex, ey,
                                              V(0), ezmj, V(0), next_x<V>(ezm, n), hxk, hyk, hz0);
            }
        }
    }
    rows.ex[k0 & 1][w][lane] = ex;
    rows.ez[k0 & 1][w][lane] = ez;
    __syncthreads();
    for (int k = k0; k < k1; ++k, c += plane) {
        cYA<real> ap = base;
        asm volatile("" : "+s"(ap));
        const YeeArgs<real> &a = *(const YeeArgs<real> *)ap;
        const Sets<real> s = sets(a, st);
        const AxC<real> yc = axc(a.ay, jj);
        const AxC<real> yh = axc(a.ay, drow ? j - 1 : jj);
        const int p = k & 1;
        const AxC<real> zc = axc(a.az, k);
        const bool top = k + 1 >= a.nz;  // E above the top plane is the zero halo
        V exk = V(0), eyk = V(0), ezk = V(0);
        V hx = V(0), hy = V(0), hz = V(0), dx = V(0), dy = V(0), dz = V(0);
        real ezn = 0, eyn = 0;
        typename Vec<real>::ids idw = 0;
        bool anyd = false;  // E-form: the lane holds a cell that keeps D
        // FP32: lane 0's new Hy, Hz of cell x0-1 depend only on old values
        // and are loaded and computed before the barrier, so those loads do
        // not wait after it.  The first wavefront's recomputed row j0-1 and
        // the D of cells that keep it come after the barrier in both forms:
        // held across it they cost 21 VGPRs, and the FP32 kernel fits 4
        // wavefronts per SIMD (127 VGPRs) only without them.
        constexpr bool early = EForm<real>::value;
        V hxj = V(0), hzj = V(0);  // new Hx, Hz of row j-1
        V eyhk = V(0);             // Ey of row j-1 at plane k+1 (first wavefront)
        real hyx = 0, hzx = 0;     // new Hy, Hz of cell x0-1 (lane 0)
        if (live) {
            if (!top) {
                exk = ld<V>(s.Eo[0] + c + plane);
                eyk = ld<V>(s.Eo[1] + c + plane);
                ezk = ld<V>(s.Eo[2] + c + plane);
            }
            hx = ld<V>(s.Ho[0] + c), hy = ld<V>(s.Ho[1] + c), hz = ld<V>(s.Ho[2] + c);
            idw = cell_ids(a, c);
            if constexpr (EForm<real>::value) anyd = any_dstate(a, idw);  // D is loaded at the store
            else dx = ldnt<V>(a.Dx + c), dy = ldnt<V>(a.Dy + c), dz = ldnt<V>(a.Dz + c);
            ezn = shfl_down1(ez[0]), eyn = shfl_down1(ey[0]);
            if (edge) {
                ezn = s.Eo[2][c + VN];
                eyn = s.Eo[1][c + VN];
            }
            if (early && drow && xhalo) h_cell_yz(a, s, yc, zc, x0 - 1, j, k, c - 1, top, ez[0], ey[0], hyx, hzx);
            if (hrow) {
                V exj, ezj;
                if (last) {
                    exj = ld<V>(s.Eo[0] + c + nxp);
                    ezj = ld<V>(s.Eo[2] + c + nxp);
                } else {
                    exj = rows.ex[p][w + 1][lane];
                    ezj = rows.ez[p][w + 1][lane];
                }
                h_cells<real, HX | HY | HZ, true>(a, s, lx, yc, zc, j, k, store, ex, ey, ez, exk, eyk, exj, ezj,
                                                  next_x<V>(ey, eyn), next_x<V>(ez, ezn), hx, hy, hz);
                if (store) {
                    stnt(hx, s.Hn[0] + c);
                    stnt(hy, s.Hn[1] + c);
                    stnt(hz, s.Hn[2] + c);
                }
            }
        }
        rows.hx[p][w][lane] = hx;
        rows.hz[p][w][lane] = hz;
        rows.ex[p ^ 1][w][lane] = exk;
        rows.ez[p ^ 1][w][lane] = ezk;
        __syncthreads();
        if (live) {
            int mid[VN], fl[VN];
            bool anyloss, allplain;
            cell_flags<real>(idw, mats, mid, fl, anyloss, allplain);
            V gx = V(0), gy = V(0), gz = V(0);  // E-form: dD
            if (drow) {
                // new Hx, Hz of row j-1: the wavefront above, or recomputed by the first
                if (!first) {
                    hxj = rows.hx[p][w - 1][lane];
                    hzj = rows.hz[p][w - 1][lane];
                } else {
                    const size_t m = c - nxp;
                    const V exh = ld<V>(s.Eo[0] + m), ezh = ld<V>(s.Eo[2] + m);
                    if (!top) eyhk = ld<V>(s.Eo[1] + m + plane);
                    real n = shfl_down1(eyh[0]);
                    if (edge) n = s.Eo[1][m + VN];
                    V hyj = V(0);
                    hxj = ld<V>(s.Ho[0] + m);
                    hzj = ld<V>(s.Ho[2] + m);
                    h_cells<real, HX | HZ, false>(a, s, lx, yh, zc, j - 1, k, false, exh, eyh, ezh, V(0), eyhk, ex,
                                                  ez, next_x<V>(eyh, n), V(0), hxj, hyj, hzj);
                }
                // new Hy, Hz of the i-1 neighbours: the previous lane's last cell;
                // lane 0 recomputes the cell before its vector
                real hyp = shfl_up1(hy[VN - 1]), hzp = shfl_up1(hz[VN - 1]);
                if (xhalo) {
                    if (early) {
                        hyp = hyx;
                        hzp = hzx;
                    } else {
                        h_cell_yz(a, s, yc, zc, x0 - 1, j, k, c - 1, top, ez[0], ey[0], hyp, hzp);
                    }
                }
                if constexpr (EForm<real>::value)  // the increment dD, from zero
                    d_cells(a, lx, yc, zc, j, k, store, mats, mid, fl, anyloss, hx, hy, hz, hxj, hzj, hxk, hyk,
                            prev_x<V>(hy, hyp), prev_x<V>(hz, hzp), ex, ey, ez, gx, gy, gz);
                else
                    d_cells(a, lx, yc, zc, j, k, store, mats, mid, fl, anyloss, hx, hy, hz, hxj, hzj, hxk, hyk,
                            prev_x<V>(hy, hyp), prev_x<V>(hz, hzp), ex, ey, ez, dx, dy, dz);
                hxk = hx;
                hyk = hy;
            }
            if constexpr (EForm<real>::value) {
                inject(a, x0, j, k, st, gx, gy, gz);
                if (anyd) dx = ldnt<V>(a.Dx + c), dy = ldnt<V>(a.Dy + c), dz = ldnt<V>(a.Dz + c);
                if (store) store_ef(a, s, c, mats, mid, fl, anyd, ex, ey, ez, dx, dy, dz, gx, gy, gz);
            } else {
                inject(a, x0, j, k, st, dx, dy, dz);
                if (store) store_de(a, s, c, mats, mid, fl, allplain, dx, dy, dz);
            }
            ex = exk;
            ey = eyk;
            ez = ezk;
            eyh = eyhk;
        }
    }
}

// Split plan, first kernel: new H of tile t (one row per wavefront).
template <typename real>
__device__ __forceinline__ void h_tile(const YeeArgs<real> &a, const int t) {
    typedef typename Vec<real>::type V;
    constexpr int VN = Vec<real>::n;
    const TileOf tl = tile_of(a, t);
    const int j = tl.j;
    const int lane = threadIdx.x;
    const int x0 = (tl.bx * 64 + lane) * VN;
    if (j >= a.ny - 1 || uniform(tl.bx * 64 * VN) >= a.nx - 1) return;
    const bool store = x0 < a.nx - 1;
    const int k0 = uniform(a.kbase + tl.bz * a.kchunk);
    const int k1 = min(k0 + a.kchunk, a.kend);
    const size_t nxp = a.nxp;
    const size_t plane = nxp * a.ny;
    const Sets<real> s = sets(a, sld(a.step, 0));
    const LaneX<real> lx = lane_x(a, x0);
    const AxC<real> yc = axc(a.ay, j);
    const bool edge = lane == 63 && x0 + VN < a.nxp;
    size_t c = ((size_t)k0 * a.ny + j) * nxp + lx.xr;
    V ex = ld<V>(s.Eo[0] + c), ey = ld<V>(s.Eo[1] + c), ez = ld<V>(s.Eo[2] + c);
    for (int k = k0; k < k1; ++k, c += plane) {
        const AxC<real> zc = axc(a.az, k);
        V exk = V(0), eyk = V(0), ezk = V(0);
        if (k + 1 < a.nz) {
            exk = ld<V>(s.Eo[0] + c + plane);
            eyk = ld<V>(s.Eo[1] + c + plane);
            ezk = ld<V>(s.Eo[2] + c + plane);
        }
        V hx = ldnt<V>(s.Ho[0] + c), hy = ldnt<V>(s.Ho[1] + c), hz = ldnt<V>(s.Ho[2] + c);
        real ezn = shfl_down1(ez[0]), eyn = shfl_down1(ey[0]);
        if (edge) {
            ezn = s.Eo[2][c + VN];
            eyn = s.Eo[1][c + VN];
        }
        const V exj = ld<V>(s.Eo[0] + c + nxp), ezj = ld<V>(s.Eo[2] + c + nxp);
        h_cells<real, HX | HY | HZ, true>(a, s, lx, yc, zc, j, k, store, ex, ey, ez, exk, eyk, exj, ezj,
                                          next_x<V>(ey, eyn), next_x<V>(ez, ezn), hx, hy, hz);
        if (store) {
            stnt(hx, s.Hn[0] + c);
            stnt(hy, s.Hn[1] + c);
            stnt(hz, s.Hn[2] + c);
        }
        ex = exk;
        ey = eyk;
        ez = ezk;
    }
}

// Split plan, second kernel: new D and E of plain cells of tile t, from the
// new H that h_update wrote.
template <typename real>
__device__ __forceinline__ void de_tile(const YeeArgs<real> &a, const int t, const MatCoef<real> *mats) {
    typedef typename Vec<real>::type V;
    constexpr int VN = Vec<real>::n;
    const TileOf tl = tile_of(a, t);
    const int j = tl.j;
    const int lane = threadIdx.x;
    const int x0 = (tl.bx * 64 + lane) * VN;
    if (j >= a.ny || uniform(tl.bx * 64 * VN) >= a.nx) return;
    const bool store = x0 < a.nx;
    const int k0 = uniform(a.kbase + tl.bz * a.kchunk);
    const int k1 = min(k0 + a.kchunk, a.kend);
    const size_t nxp = a.nxp;
    const size_t plane = nxp * a.ny;
    const int64_t st = sld(a.step, 0);
    const Sets<real> s = sets(a, st);
    const real *const *Hn = s.Hn;
    const LaneX<real> lx = lane_x(a, x0);
    const bool drow = j >= 1 && j <= a.ny - 2;
    const AxC<real> yc = axc(a.ay, j);
    const bool edge = lane == 0 && x0 >= 1;  // reads the new H of the cell before its vector
    size_t c = ((size_t)k0 * a.ny + j) * nxp + lx.xr;
    V hxk = V(0), hyk = V(0);
    if (drow && k0 > 0) {
        hxk = ld<V>(Hn[0] + c - plane);
        hyk = ld<V>(Hn[1] + c - plane);
    }
    for (int k = k0; k < k1; ++k, c += plane) {
        const AxC<real> zc = axc(a.az, k);
        constexpr bool EF = EForm<real>::value;
        int mid[VN], fl[VN];
        bool anyloss, allplain;
        cell_mats(a, c, mats, mid, fl, anyloss, allplain);
        // D-form: D, advanced in place.  E-form: D only for lanes with a cell
        // that keeps it, the increment dD in g, and E^n of every cell.
        const bool anyd = EF && any_dstate(a, *reinterpret_cast<const typename Vec<real>::ids *>(a.medium + c));
        V dx = V(0), dy = V(0), dz = V(0), gx = V(0), gy = V(0), gz = V(0);
        if (!EF || anyd) dx = ldnt<V>(a.Dx + c), dy = ldnt<V>(a.Dy + c), dz = ldnt<V>(a.Dz + c);
        V ex = V(0), ey = V(0), ez = V(0);
        if (EF || (drow && anyloss)) {
            ex = ld<V>(s.Eo[0] + c);
            ey = ld<V>(s.Eo[1] + c);
            ez = ld<V>(s.Eo[2] + c);
        }
        V &ux = EF ? gx : dx, &uy = EF ? gy : dy, &uz = EF ? gz : dz;  // what the curl advances
        if (drow) {
            const V hx = ld<V>(Hn[0] + c), hy = ld<V>(Hn[1] + c), hz = ld<V>(Hn[2] + c);
            const V hxj = ld<V>(Hn[0] + c - nxp), hzj = ld<V>(Hn[2] + c - nxp);
            real hyp = shfl_up1(hy[VN - 1]), hzp = shfl_up1(hz[VN - 1]);
            if (edge) {
                hyp = Hn[1][c - 1];
                hzp = Hn[2][c - 1];
            }
            d_cells(a, lx, yc, zc, j, k, store, mats, mid, fl, anyloss, hx, hy, hz, hxj, hzj, hxk, hyk,
                    prev_x<V>(hy, hyp), prev_x<V>(hz, hzp), ex, ey, ez, ux, uy, uz);
            hxk = hx;
            hyk = hy;
        }
        inject(a, x0, j, k, st, ux, uy, uz);
        if (store) {
            if constexpr (EF) store_ef(a, s, c, mats, mid, fl, anyd, ex, ey, ez, dx, dy, dz, gx, gy, gz);
            else store_de(a, s, c, mats, mid, fl, allplain, dx, dy, dz);
        }
    }
}

// Material table of a domain into LDS (all threads of the work-group).
template <typename real>
__device__ __forceinline__ MatCoef<real> *stage_mats(const YeeArgs<real> &a) {
    extern __shared__ unsigned char lds_raw[];
    const int nw = a.nmat * (int)(sizeof(MatCoef<real>) / 4);
    const uint32_t *src = reinterpret_cast<const uint32_t *>(a.mats);
    uint32_t *dst = reinterpret_cast<uint32_t *>(lds_raw);
    for (int w = threadIdx.y * 64 + threadIdx.x; w < nw; w += 64 * (int)blockDim.y) dst[w] = src[w];
    __syncthreads();
    return reinterpret_cast<MatCoef<real> *>(lds_raw);
}

template <typename real, int ROWS>
__global__ __launch_bounds__(64 * ROWS) void yee_update(YeeArgs<real> a) {
    const MatCoef<real> *mats = stage_mats(a);
    const int t = band_tile(blockIdx.x, a.nblocks, a.nbx);
    if (t < 0) return;
    yee_tile<real, ROWS>(a, (cYA<real>)__builtin_amdgcn_kernarg_segment_ptr(), t, mats);
}

template <typename real>
__global__ __launch_bounds__(256) void h_update(YeeArgs<real> a) {
    const int t = band_tile(blockIdx.x, a.nblocks, a.nbx);
    if (t < 0) return;
    h_tile<real>(a, t);
}

template <typename real>
__global__ __launch_bounds__(256) void de_update(YeeArgs<real> a) {
    const MatCoef<real> *mats = stage_mats(a);
    const int t = band_tile(blockIdx.x, a.nblocks, a.nbx);
    if (t < 0) return;
    de_tile<real>(a, t, mats);
}

// Batched launches: the tiles of every domain of a batch in one grid.  The
// XCD renumbering runs over the concatenated tiles.

// One-pass kernel over the tiles of every domain of a batch (8-row
// work-groups, as alone).  Each domain's argument block in the batch table
// is read-only for the launch, so the plane loop re-reads it there.
template <typename real>
__global__ __launch_bounds__(64 * kBatchRows) void yee_update_batch(BatchLaunch<YeeArgs<real>> b) {
    const int g = xcd_tile(blockIdx.x, b.total, b.total_pad);
    if (g < 0) return;  // whole work-group
    const int rid = (int)sld(b.map, g);
    const YeeArgs<real> a = batch_args(b, rid);
    const MatCoef<real> *mats = stage_mats(a);
    yee_tile<real, kBatchRows>(a, (cYA<real>)(b.args + rid), g - sld(b.start, rid), mats);
}

template <typename real>
__global__ __launch_bounds__(256) void h_update_batch(BatchLaunch<YeeArgs<real>> b) {
    const int g = xcd_tile(blockIdx.x, b.total, b.total_pad);
    if (g < 0) return;
    const int rid = (int)sld(b.map, g);
    const YeeArgs<real> a = batch_args(b, rid);
    h_tile<real>(a, g - sld(b.start, rid));
}

template <typename real>
__global__ __launch_bounds__(256) void de_update_batch(BatchLaunch<YeeArgs<real>> b) {
    const int g = xcd_tile(blockIdx.x, b.total, b.total_pad);
    if (g < 0) return;  // whole work-group
    const int rid = (int)sld(b.map, g);
    const YeeArgs<real> a = batch_args(b, rid);
    const MatCoef<real> *mats = stage_mats(a);
    de_tile<real>(a, g - sld(b.start, rid), mats);
}

template <typename real>
void launch_yee(const YeeArgs<real> &a, hipStream_t s) {
    const size_t lds = (size_t)a.nmat * sizeof(MatCoef<real>);
    switch (a.rows) {
        case 6: hipLaunchKernelGGL((yee_update<real, 6>), dim3(a.nblocks_pad), dim3(64, 6), lds, s, a); return;
        case 8: hipLaunchKernelGGL((yee_update<real, 8>), dim3(a.nblocks_pad), dim3(64, 8), lds, s, a); return;
        case 12: hipLaunchKernelGGL((yee_update<real, 12>), dim3(a.nblocks_pad), dim3(64, 12), lds, s, a); return;
        case 16: hipLaunchKernelGGL((yee_update<real, 16>), dim3(a.nblocks_pad), dim3(64, 16), lds, s, a); return;
        default: hipLaunchKernelGGL((yee_update<real, 4>), dim3(a.nblocks_pad), dim3(64, 4), lds, s, a);
    }
}

template <typename real>
void launch_h(const YeeArgs<real> &a, hipStream_t s) {
    hipLaunchKernelGGL(h_update<real>, dim3(a.nblocks_pad), dim3(64, kBlockY), 0, s, a);
}

template <typename real>
void launch_de(const YeeArgs<real> &a, hipStream_t s) {
    const size_t lds = (size_t)a.nmat * sizeof(MatCoef<real>);
    hipLaunchKernelGGL(de_update<real>, dim3(a.nblocks_pad), dim3(64, kBlockY), lds, s, a);
}

template <typename real>
void launch_yee_batch(const BatchLaunch<YeeArgs<real>> &b, int max_nmat, hipStream_t s) {
    const size_t lds = (size_t)max_nmat * sizeof(MatCoef<real>);
    if (b.total > 0)
        hipLaunchKernelGGL((yee_update_batch<real>), dim3(b.total_pad), dim3(64, kBatchRows), lds, s, b);
}

template <typename real>
void launch_h_batch(const BatchLaunch<YeeArgs<real>> &b, hipStream_t s) {
    if (b.total > 0) hipLaunchKernelGGL(h_update_batch<real>, dim3(b.total_pad), dim3(64, kBlockY), 0, s, b);
}

template <typename real>
void launch_de_batch(const BatchLaunch<YeeArgs<real>> &b, int max_nmat, hipStream_t s) {
    const size_t lds = (size_t)max_nmat * sizeof(MatCoef<real>);
    if (b.total > 0) hipLaunchKernelGGL(de_update_batch<real>, dim3(b.total_pad), dim3(64, kBlockY), lds, s, b);
}

template void launch_yee<double>(const YeeArgs<double> &, hipStream_t);
template void launch_yee<float>(const YeeArgs<float> &, hipStream_t);
template void launch_h<double>(const YeeArgs<double> &, hipStream_t);
template void launch_h<float>(const YeeArgs<float> &, hipStream_t);
template void launch_de<double>(const YeeArgs<double> &, hipStream_t);
template void launch_de<float>(const YeeArgs<float> &, hipStream_t);
template void launch_yee_batch<double>(const BatchLaunch<YeeArgs<double>> &, int, hipStream_t);
template void launch_yee_batch<float>(const BatchLaunch<YeeArgs<float>> &, int, hipStream_t);
template void launch_h_batch<double>(const BatchLaunch<YeeArgs<double>> &, hipStream_t);
template void launch_h_batch<float>(const BatchLaunch<YeeArgs<float>> &, hipStream_t);
template void launch_de_batch<double>(const BatchLaunch<YeeArgs<double>> &, int, hipStream_t);
template void launch_de_batch<float>(const BatchLaunch<YeeArgs<float>> &, int, hipStream_t);

}  // namespace phpfdtd
