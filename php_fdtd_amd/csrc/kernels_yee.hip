// kernels_yee.hip - H update and fused D/E update for gfx950.
//
// Work decomposition: every lane owns a 16-byte vector of consecutive x cells
// (4 cells in FP32, 2 in FP64), so a wavefront covers a 1 KiB x segment of a
// row and every field stream is read and written in 16-byte lane accesses,
// the width at which the HBM3E streams run fastest on this part.  A
// wavefront owns R consecutive y rows and marches up a short chunk of z
// planes:
//   * z neighbour: E at k+1 (H kernel) / H at k-1 (D kernel) is carried in
//     registers from the previous plane;
//   * y neighbour: rows r+1 (H) / r-1 (D) come from registers, one halo row
//     per R rows is loaded from the L2 / Infinity Cache copy;
//   * x neighbour: the adjacent cell of the lane's vector, or element 0 (H) /
//     the last element (D) of the neighbouring lane via a cross-lane shuffle;
//     only the wavefront's edge lane loads one extra cell.
// j and k are wavefront-uniform.  Hardware places work-group b on XCD
// (b mod 8); xcd_tile() renumbers the groups so that each XCD owns one
// contiguous range of tiles and neighbouring tiles share its L2.
//
// Reference semantics (fdtd.c):
//   H: i < nx-1, j < ny-1, all k; E above the top plane is 0 (3123-3231).
//   D: 1 <= i <= nx-2, 1 <= j <= ny-2, all k; H below k=0 is 0 (3238-3356).
//   CPML psi is lagged: the old psi enters the field, then psi is advanced.
//   psi is stored pre-multiplied by dt.
//   D loss uses E^n in the lab frame with unrotated eps_im (3349-3353).
//   Source: D[comp] += value*dt after the curl and loss (3363-3396).
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

template <typename V>
__device__ __forceinline__ V vld(const V *p) {
    return __builtin_nontemporal_load(p);
}
template <typename V>
__device__ __forceinline__ void vst(V v, V *p) {
    __builtin_nontemporal_store(v, p);
}

template <typename real>
__device__ __forceinline__ real shfl_down1(real v) {
    return __shfl_down(v, 1, 64);
}
template <typename real>
__device__ __forceinline__ real shfl_up1(real v) {
    return __shfl_up(v, 1, 64);
}

// H update of tile t of one domain.
template <typename real, int R>
__device__ __forceinline__ void h_tile(const YeeArgs<real> &a, const int t) {
    constexpr int VN = Vec<real>::n;
    typedef typename Vec<real>::type V;
    const int bx = t % a.nbx;
    const int rest = t / a.nbx;
    const int by = rest % a.nby;
    const int bz = rest / a.nby;
    const int lane = threadIdx.x;
    const int x0 = (bx * 64 + lane) * VN;  // first cell of this lane
    const int j0 = uniform((by * kBlockY + threadIdx.y) * R);
    const int nr = min(R, a.ny - 1 - j0);  // H rows j < ny-1 of this wavefront
    if (nr <= 0 || uniform(bx * 64 * VN) >= a.nx - 1) return;
    const bool lane_on = x0 < a.nx - 1;  // lanes past the row end still load and shuffle
    const int xr = x0 < a.nxp ? x0 : 0;  // in-row offset (lanes past the padded row read cell 0)
    const int k0 = uniform(a.kbase + bz * a.kchunk);
    const int k1 = min(k0 + a.kchunk, a.kend);
    const size_t nxp = a.nxp;
    const size_t plane = nxp * a.ny;
    const int L2 = 2 * a.L;
    const bool edge = lane == 63;
    const bool edge_in = x0 + VN < a.nxp;  // the cell after this lane's vector exists

    // per-cell x coefficients (tables padded to nxp)
    const V khx = *reinterpret_cast<const V *>(a.ax.kh + xr);
    int px[VN];
    bool vx = true;  // x-PML lane whose VN cells are consecutive slots of one aligned slab vector
    const int sl0 = a.ax.slot[xr];
#pragma unroll
    for (int e = 0; e < VN; ++e) {
        const int sl = a.ax.slot[xr + e];
        px[e] = x0 + e < a.nx - 1 ? sl : -1;
        vx = vx && sl == sl0 + e;
    }
    // such lanes take the vector psi path, the others go cell by cell; FP32
    // only (2.9% off the FP32 step, neutral in FP64 where it costs registers)
    vx = vx && sizeof(real) == 4 && L2 % VN == 0 && sl0 >= 0 && sl0 % VN == 0;
    real khy[R], byr[R], chr_[R];
    int py[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int j = min(j0 + r, a.ny - 1);
        khy[r] = sld(a.ay.kh, j);
        py[r] = sld(a.ay.slot, j);
        byr[r] = sld(a.ay.b, j);
        chr_[r] = sld(a.ay.ch, j);
    }
    bool anypx = false;
#pragma unroll
    for (int e = 0; e < VN; ++e) anypx |= px[e] >= 0;

    size_t c0 = ((size_t)k0 * a.ny + j0) * nxp + xr;
    V ex[R + 1], ey[R];
#pragma unroll
    for (int r = 0; r <= R; ++r)
        if (r <= nr) ex[r] = *reinterpret_cast<const V *>(a.Ex + c0 + r * nxp);
#pragma unroll
    for (int r = 0; r < R; ++r)
        if (r < nr) ey[r] = *reinterpret_cast<const V *>(a.Ey + c0 + r * nxp);

    for (int k = k0; k < k1; ++k, c0 += plane) {
        V exn[R + 1], eyn[R], ez[R + 1];
        const bool top = k + 1 >= a.nz;  // E above the top plane is the zero halo
#pragma unroll
        for (int r = 0; r <= R; ++r)
            if (r <= nr) {
                exn[r] = top ? V(0) : *reinterpret_cast<const V *>(a.Ex + c0 + plane + r * nxp);
                ez[r] = *reinterpret_cast<const V *>(a.Ez + c0 + r * nxp);
            }
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (r < nr) eyn[r] = top ? V(0) : *reinterpret_cast<const V *>(a.Ey + c0 + plane + r * nxp);
        const real khz = sld(a.az.kh, k);
        const int pz = sld(a.az.slot, k);
        const real bzv = sld(a.az.b, k), chz = sld(a.az.ch, k);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (r >= nr) continue;
            const size_t c = c0 + r * nxp;
            const int j = j0 + r;
            // x neighbours: element e+1 of the vector, element 0 of the next lane
            real ez_n = shfl_down1(ez[r][0]), ey_n = shfl_down1(ey[r][0]);
            if (edge && edge_in) {
                ez_n = a.Ez[c + VN];
                ey_n = a.Ey[c + VN];
            }
            V ezip, eyip;
#pragma unroll
            for (int e = 0; e < VN - 1; ++e) {
                ezip[e] = ez[r][e + 1];
                eyip[e] = ey[r][e + 1];
            }
            ezip[VN - 1] = ez_n;
            eyip[VN - 1] = ey_n;
            const V dyEz = ez[r + 1] - ez[r], dzEy = eyn[r] - ey[r];
            const V dxEz = ezip - ez[r], dzEx = exn[r] - ex[r];
            const V dxEy = eyip - ey[r], dyEx = ex[r + 1] - ex[r];
            V hx = vld(reinterpret_cast<const V *>(a.Hx + c));
            V hy = vld(reinterpret_cast<const V *>(a.Hy + c));
            V hz = vld(reinterpret_cast<const V *>(a.Hz + c));
            V nhx = hx + khz * dzEy - khy[r] * dyEz;
            V nhy = hy + khx * dxEz - khz * dzEx;
            V nhz = hz + khy[r] * dyEx - khx * dxEy;
            if (vx) {
                V *sxyp = reinterpret_cast<V *>(a.pHxy + ((size_t)k * a.ny + j) * L2 + sl0);
                V *sxzp = reinterpret_cast<V *>(a.pHxz + ((size_t)k * a.ny + j) * L2 + sl0);
                const V sxy = *sxyp, sxz = *sxzp;
                const V bxv = *reinterpret_cast<const V *>(a.ax.b + xr);
                const V chxv = *reinterpret_cast<const V *>(a.ax.ch + xr);
                nhy += sxy;
                nhz += sxz;
                V nsxy = bxv * sxy + chxv * dxEz, nsxz = bxv * sxz - chxv * dxEy;
#pragma unroll
                for (int e = 0; e < VN; ++e)
                    if (x0 + e >= a.nx - 1) { nsxy[e] = sxy[e]; nsxz[e] = sxz[e]; }
                if (lane_on) { *sxyp = nsxy; *sxzp = nsxz; }
            } else if (anypx) {
#pragma unroll
                for (int e = 0; e < VN; ++e) {
                    if (px[e] < 0) continue;
                    const int i = xr + e;
                    const size_t q = ((size_t)k * a.ny + j) * L2 + px[e];
                    const real sxy = a.pHxy[q], sxz = a.pHxz[q];
                    nhy[e] += sxy;
                    nhz[e] += sxz;
                    a.pHxy[q] = a.ax.b[i] * sxy + a.ax.ch[i] * dxEz[e];
                    a.pHxz[q] = a.ax.b[i] * sxz - a.ax.ch[i] * dxEy[e];
                }
            }
            if (py[r] >= 0) {
                V *syxp = reinterpret_cast<V *>(a.pHyx + ((size_t)k * L2 + py[r]) * nxp + xr);
                V *syzp = reinterpret_cast<V *>(a.pHyz + ((size_t)k * L2 + py[r]) * nxp + xr);
                const real byv = byr[r], chy = chr_[r];
                const V syx = *syxp, syz = *syzp;
                nhx += syx;
                nhz += syz;
                V nsyx = byv * syx - chy * dyEz, nsyz = byv * syz + chy * dyEx;
#pragma unroll
                for (int e = 0; e < VN; ++e)
                    if (x0 + e >= a.nx - 1) { nsyx[e] = syx[e]; nsyz[e] = syz[e]; }
                if (lane_on) { *syxp = nsyx; *syzp = nsyz; }
            }
            if (pz >= 0) {
                V *szxp = reinterpret_cast<V *>(a.pHzx + ((size_t)pz * a.ny + j) * nxp + xr);
                V *szyp = reinterpret_cast<V *>(a.pHzy + ((size_t)pz * a.ny + j) * nxp + xr);
                const V szx = *szxp, szy = *szyp;
                nhx += szx;
                nhy += szy;
                V nszx = bzv * szx + chz * dzEy, nszy = bzv * szy - chz * dzEx;
#pragma unroll
                for (int e = 0; e < VN; ++e)
                    if (x0 + e >= a.nx - 1) { nszx[e] = szx[e]; nszy[e] = szy[e]; }
                if (lane_on) { *szxp = nszx; *szyp = nszy; }
            }
            // cells i >= nx-1 keep their value (never updated by the reference)
#pragma unroll
            for (int e = 0; e < VN; ++e)
                if (x0 + e >= a.nx - 1) {
                    nhx[e] = hx[e];
                    nhy[e] = hy[e];
                    nhz[e] = hz[e];
                }
            if (lane_on) {
                vst(nhx, reinterpret_cast<V *>(a.Hx + c));
                vst(nhy, reinterpret_cast<V *>(a.Hy + c));
                vst(nhz, reinterpret_cast<V *>(a.Hz + c));
            }
        }
#pragma unroll
        for (int r = 0; r <= R; ++r) ex[r] = exn[r];
#pragma unroll
        for (int r = 0; r < R; ++r) ey[r] = eyn[r];
    }
}

// Material table of a domain into LDS (all threads of the work-group).
template <typename real>
__device__ __forceinline__ MatCoef<real> *stage_mats(const YeeArgs<real> &a) {
    extern __shared__ unsigned char lds_raw[];
    const int nw = a.nmat * (int)(sizeof(MatCoef<real>) / 4);
    const uint32_t *src = reinterpret_cast<const uint32_t *>(a.mats);
    uint32_t *dst = reinterpret_cast<uint32_t *>(lds_raw);
    for (int w = threadIdx.y * 64 + threadIdx.x; w < nw; w += 256) dst[w] = src[w];
    __syncthreads();
    return reinterpret_cast<MatCoef<real> *>(lds_raw);
}

// D update and E of plain cells for tile t of one domain; mats is the LDS
// copy of the domain's material table.
template <typename real, int R>
__device__ __forceinline__ void de_tile(const YeeArgs<real> &a, const int t, const MatCoef<real> *mats) {
    constexpr int VN = Vec<real>::n;
    typedef typename Vec<real>::type V;
    typedef typename Vec<real>::ids IDS;
    const int bx = t % a.nbx;
    const int rest = t / a.nbx;
    const int by = rest % a.nby;
    const int bz = rest / a.nby;
    const int lane = threadIdx.x;
    const int x0 = (bx * 64 + lane) * VN;
    const int j0 = uniform((by * kBlockY + threadIdx.y) * R);
    const int nr = min(R, a.ny - j0);  // E/D rows of this wavefront
    if (nr <= 0 || uniform(bx * 64 * VN) >= a.nx) return;
    const bool lane_on = x0 < a.nx;
    const int xr = x0 < a.nxp ? x0 : 0;
    const int k0 = uniform(a.kbase + bz * a.kchunk);
    const int k1 = min(k0 + a.kchunk, a.kend);
    const size_t nxp = a.nxp;
    const size_t plane = nxp * a.ny;
    const int L2 = 2 * a.L;
    const bool edge = lane == 0 && x0 >= 1;  // needs the cell before its vector from memory

    bool inx[VN];
    int px[VN];
#pragma unroll
    for (int e = 0; e < VN; ++e) {
        const int i = x0 + e;
        inx[e] = i >= 1 && i <= a.nx - 2;
        px[e] = a.ax.slot[xr + e];
    }
    const V kdx = *reinterpret_cast<const V *>(a.ax.kd + xr);
    bool anypx = false;
#pragma unroll
    for (int e = 0; e < VN; ++e) anypx |= px[e] >= 0 && inx[e];
    // vector psi path as in h_tile (FP32)
    bool vx = sizeof(real) == 4 && L2 % VN == 0 && px[0] >= 0 && px[0] % VN == 0;
#pragma unroll
    for (int e = 1; e < VN; ++e) vx = vx && px[e] == px[0] + e;
    bool iny[R];
    real kdy[R], byr[R], cdr[R];
    int py[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int j = j0 + r;
        iny[r] = j >= 1 && j <= a.ny - 2 && r < nr;
        const int jc = min(j, a.ny - 1);
        kdy[r] = sld(a.ay.kd, jc);
        py[r] = sld(a.ay.slot, jc);
        byr[r] = sld(a.ay.b, jc);
        cdr[r] = sld(a.ay.cd, jc);
    }
    const bool halo = j0 >= 1;

    const int64_t st = a.nsrc > 0 ? sld(a.step, 0) : 0;

    size_t c0 = ((size_t)k0 * a.ny + j0) * nxp + xr;
    V hxm[R], hym[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        hxm[r] = V(0);
        hym[r] = V(0);
        if (r < nr && k0 > 0) {
            hxm[r] = *reinterpret_cast<const V *>(a.Hx + c0 - plane + r * nxp);
            hym[r] = *reinterpret_cast<const V *>(a.Hy + c0 - plane + r * nxp);
        }
    }
    for (int k = k0; k < k1; ++k, c0 += plane) {
        V hx[R + 1], hz[R + 1], hy[R];
        hx[0] = V(0);
        hz[0] = V(0);
        if (halo) {
            hx[0] = *reinterpret_cast<const V *>(a.Hx + c0 - nxp);
            hz[0] = *reinterpret_cast<const V *>(a.Hz + c0 - nxp);
        }
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (r < nr) {
                hx[r + 1] = *reinterpret_cast<const V *>(a.Hx + c0 + r * nxp);
                hz[r + 1] = *reinterpret_cast<const V *>(a.Hz + c0 + r * nxp);
                hy[r] = *reinterpret_cast<const V *>(a.Hy + c0 + r * nxp);
            }
        const real kdz = sld(a.az.kd, k);
        const int pz = sld(a.az.slot, k);
        const real bzv = sld(a.az.b, k), cdz = sld(a.az.cd, k);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (r >= nr) continue;
            const size_t c = c0 + r * nxp;
            const int j = j0 + r;
            V dx = vld(reinterpret_cast<const V *>(a.Dx + c));
            V dy = vld(reinterpret_cast<const V *>(a.Dy + c));
            V dz = vld(reinterpret_cast<const V *>(a.Dz + c));
            const IDS idw = *reinterpret_cast<const IDS *>(a.medium + c);
            int mid[VN], fl[VN], anyloss = 0, allplain = 1;
#pragma unroll
            for (int e = 0; e < VN; ++e) {
                mid[e] = (idw >> (8 * e)) & 0xff;
                fl[e] = mats[mid[e]].flags;
                anyloss |= fl[e] & MF_LOSS;
                allplain &= !(fl[e] & MF_SPECIAL);
            }
            // x neighbours at i-1: element e-1, or the last element of the previous lane
            real hy_p = shfl_up1(hy[r][VN - 1]), hz_p = shfl_up1(hz[r + 1][VN - 1]);
            if (edge) {
                hy_p = a.Hy[c - 1];
                hz_p = a.Hz[c - 1];
            }
            V hyim, hzim;
            hyim[0] = hy_p;
            hzim[0] = hz_p;
#pragma unroll
            for (int e = 1; e < VN; ++e) {
                hyim[e] = hy[r][e - 1];
                hzim[e] = hz[r + 1][e - 1];
            }
            if (iny[r]) {
                const V dyHz = hz[r + 1] - hz[r], dzHy = hy[r] - hym[r];
                const V dzHx = hx[r + 1] - hxm[r], dxHz = hz[r + 1] - hzim;
                const V dxHy = hy[r] - hyim, dyHx = hx[r + 1] - hx[r];
                V ndx = dx + kdy[r] * dyHz - kdz * dzHy;
                V ndy = dy + kdz * dzHx - kdx * dxHz;
                V ndz = dz + kdx * dxHy - kdy[r] * dyHx;
                if (vx) {
                    V *sxyp = reinterpret_cast<V *>(a.pDxy + ((size_t)k * a.ny + j) * L2 + px[0]);
                    V *sxzp = reinterpret_cast<V *>(a.pDxz + ((size_t)k * a.ny + j) * L2 + px[0]);
                    const V sxy = *sxyp, sxz = *sxzp;
                    const V bxv = *reinterpret_cast<const V *>(a.ax.b + xr);
                    const V cdxv = *reinterpret_cast<const V *>(a.ax.cd + xr);
                    ndy += sxy;
                    ndz += sxz;
                    V nsxy = bxv * sxy - cdxv * dxHz, nsxz = bxv * sxz + cdxv * dxHy;
#pragma unroll
                    for (int e = 0; e < VN; ++e)
                        if (!inx[e]) { nsxy[e] = sxy[e]; nsxz[e] = sxz[e]; }
                    if (lane_on) { *sxyp = nsxy; *sxzp = nsxz; }
                } else if (anypx) {
#pragma unroll
                    for (int e = 0; e < VN; ++e) {
                        if (px[e] < 0 || !inx[e]) continue;
                        const int i = xr + e;
                        const size_t q = ((size_t)k * a.ny + j) * L2 + px[e];
                        const real sxy = a.pDxy[q], sxz = a.pDxz[q];
                        ndy[e] += sxy;
                        ndz[e] += sxz;
                        a.pDxy[q] = a.ax.b[i] * sxy - a.ax.cd[i] * dxHz[e];
                        a.pDxz[q] = a.ax.b[i] * sxz + a.ax.cd[i] * dxHy[e];
                    }
                }
                if (py[r] >= 0) {
                    V *syxp = reinterpret_cast<V *>(a.pDyx + ((size_t)k * L2 + py[r]) * nxp + xr);
                    V *syzp = reinterpret_cast<V *>(a.pDyz + ((size_t)k * L2 + py[r]) * nxp + xr);
                    const real byv = byr[r], cdy = cdr[r];
                    const V syx = *syxp, syz = *syzp;
                    ndx += syx;
                    ndz += syz;
                    // psi of cells outside the D loop (i = 0, nx-1) is never advanced
                    V nsyx = byv * syx + cdy * dyHz, nsyz = byv * syz - cdy * dyHx;
#pragma unroll
                    for (int e = 0; e < VN; ++e)
                        if (!inx[e]) { nsyx[e] = syx[e]; nsyz[e] = syz[e]; }
                    if (lane_on) { *syxp = nsyx; *syzp = nsyz; }
                }
                if (pz >= 0) {
                    V *szxp = reinterpret_cast<V *>(a.pDzx + ((size_t)pz * a.ny + j) * nxp + xr);
                    V *szyp = reinterpret_cast<V *>(a.pDzy + ((size_t)pz * a.ny + j) * nxp + xr);
                    const V szx = *szxp, szy = *szyp;
                    ndx += szx;
                    ndy += szy;
                    V nszx = bzv * szx - cdz * dzHy, nszy = bzv * szy + cdz * dzHx;
#pragma unroll
                    for (int e = 0; e < VN; ++e)
                        if (!inx[e]) { nszx[e] = szx[e]; nszy[e] = szy[e]; }
                    if (lane_on) { *szxp = nszx; *szyp = nszy; }
                }
                if (anyloss) {
                    const V ex = *reinterpret_cast<const V *>(a.Ex + c);
                    const V ey = *reinterpret_cast<const V *>(a.Ey + c);
                    const V ez = *reinterpret_cast<const V *>(a.Ez + c);
#pragma unroll
                    for (int e = 0; e < VN; ++e) {
                        if (!(fl[e] & MF_LOSS)) continue;
                        const real *Lk = mats[mid[e]].Lk;
                        ndx[e] -= Lk[0] * ex[e] + Lk[1] * ey[e] + Lk[2] * ez[e];
                        ndy[e] -= Lk[3] * ex[e] + Lk[4] * ey[e] + Lk[5] * ez[e];
                        ndz[e] -= Lk[6] * ex[e] + Lk[7] * ey[e] + Lk[8] * ez[e];
                    }
                }
#pragma unroll
                for (int e = 0; e < VN; ++e)
                    if (inx[e]) {
                        dx[e] = ndx[e];
                        dy[e] = ndy[e];
                        dz[e] = ndz[e];
                    }
            }
            for (int s = 0; s < a.nsrc; ++s) {
                const int si = sld(a.src, 4 * s), sj = sld(a.src, 4 * s + 1), sk = sld(a.src, 4 * s + 2),
                          sc = sld(a.src, 4 * s + 3);
                if (sk == k && sj == j) {
                    const int e = si - x0;
                    if (e >= 0 && e < VN) {
                        const real v = sld(a.src_val, st * a.nsrc + s);
#pragma unroll
                        for (int q = 0; q < VN; ++q)
                            if (q == e) {
                                if (sc == 0) dx[q] += v;
                                else if (sc == 1) dy[q] += v;
                                else dz[q] += v;
                            }
                    }
                }
            }
            if (lane_on) {
                vst(dx, reinterpret_cast<V *>(a.Dx + c));
                vst(dy, reinterpret_cast<V *>(a.Dy + c));
                vst(dz, reinterpret_cast<V *>(a.Dz + c));
                if (allplain) {
                    V nex, ney, nez;
#pragma unroll
                    for (int e = 0; e < VN; ++e) {
                        const real *M = mats[mid[e]].M;
                        nex[e] = M[0] * dx[e] + M[1] * dy[e] + M[2] * dz[e];
                        ney[e] = M[3] * dx[e] + M[4] * dy[e] + M[5] * dz[e];
                        nez[e] = M[6] * dx[e] + M[7] * dy[e] + M[8] * dz[e];
                    }
                    vst(nex, reinterpret_cast<V *>(a.Ex + c));
                    vst(ney, reinterpret_cast<V *>(a.Ey + c));
                    vst(nez, reinterpret_cast<V *>(a.Ez + c));
                } else {
#pragma unroll
                    for (int e = 0; e < VN; ++e) {
                        if (fl[e] & MF_SPECIAL) continue;
                        const real *M = mats[mid[e]].M;
                        a.Ex[c + e] = M[0] * dx[e] + M[1] * dy[e] + M[2] * dz[e];
                        a.Ey[c + e] = M[3] * dx[e] + M[4] * dy[e] + M[5] * dz[e];
                        a.Ez[c + e] = M[6] * dx[e] + M[7] * dy[e] + M[8] * dz[e];
                    }
                }
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            hxm[r] = hx[r + 1];
            hym[r] = hy[r];
        }
    }
}

template <typename real, int R>
__global__ __launch_bounds__(256) void h_update(YeeArgs<real> a) {
    const int t = xcd_tile(blockIdx.x, a.nblocks, a.nblocks_pad);
    if (t < 0) return;
    h_tile<real, R>(a, t);
}

template <typename real, int R>
__global__ __launch_bounds__(256) void de_update(YeeArgs<real> a) {
    const MatCoef<real> *mats = stage_mats(a);
    const int t = xcd_tile(blockIdx.x, a.nblocks, a.nblocks_pad);
    if (t < 0) return;
    de_tile<real, R>(a, t, mats);
}

// Batched launches: the tiles of every domain of a batch in one grid.  The
// XCD renumbering runs over the concatenated tiles.

template <typename real, int R>
__global__ __launch_bounds__(256) void h_update_batch(BatchLaunch<YeeArgs<real>> b) {
    const int g = xcd_tile(blockIdx.x, b.total, b.total_pad);
    if (g < 0) return;
    const int rid = (int)sld(b.map, g);
    const YeeArgs<real> a = batch_args(b, rid);
    h_tile<real, R>(a, g - sld(b.start, rid));
}

template <typename real, int R>
__global__ __launch_bounds__(256) void de_update_batch(BatchLaunch<YeeArgs<real>> b) {
    const int g = xcd_tile(blockIdx.x, b.total, b.total_pad);
    if (g < 0) return;  // whole work-group
    const int rid = (int)sld(b.map, g);
    const YeeArgs<real> a = batch_args(b, rid);
    const MatCoef<real> *mats = stage_mats(a);
    de_tile<real, R>(a, g - sld(b.start, rid), mats);
}

template <typename real, int R>
static void launch_h_r(const YeeArgs<real> &a, hipStream_t s) {
    hipLaunchKernelGGL((h_update<real, R>), dim3(a.nblocks_pad), dim3(64, kBlockY), 0, s, a);
}

template <typename real, int R>
static void launch_de_r(const YeeArgs<real> &a, hipStream_t s) {
    const size_t lds = (size_t)a.nmat * sizeof(MatCoef<real>);
    hipLaunchKernelGGL((de_update<real, R>), dim3(a.nblocks_pad), dim3(64, kBlockY), lds, s, a);
}

template <typename real>
void launch_h(const YeeArgs<real> &a, hipStream_t s) {
    switch (a.rows) {
        case 1: launch_h_r<real, 1>(a, s); break;
        case 2: launch_h_r<real, 2>(a, s); break;
        default: launch_h_r<real, 4>(a, s); break;
    }
}

template <typename real>
void launch_de(const YeeArgs<real> &a, hipStream_t s) {
    switch (a.rows) {
        case 1: launch_de_r<real, 1>(a, s); break;
        case 2: launch_de_r<real, 2>(a, s); break;
        default: launch_de_r<real, 4>(a, s); break;
    }
}

// Batched H and D/E over the domains of a batch (one y row per wavefront).
template <typename real>
void launch_h_batch(const BatchLaunch<YeeArgs<real>> &b, hipStream_t s) {
    if (b.total > 0) hipLaunchKernelGGL((h_update_batch<real, 1>), dim3(b.total_pad), dim3(64, kBlockY), 0, s, b);
}

template <typename real>
void launch_de_batch(const BatchLaunch<YeeArgs<real>> &b, int max_nmat, hipStream_t s) {
    const size_t lds = (size_t)max_nmat * sizeof(MatCoef<real>);
    if (b.total > 0)
        hipLaunchKernelGGL((de_update_batch<real, 1>), dim3(b.total_pad), dim3(64, kBlockY), lds, s, b);
}

template void launch_h<double>(const YeeArgs<double> &, hipStream_t);
template void launch_h<float>(const YeeArgs<float> &, hipStream_t);
template void launch_de<double>(const YeeArgs<double> &, hipStream_t);
template void launch_de<float>(const YeeArgs<float> &, hipStream_t);
template void launch_h_batch<double>(const BatchLaunch<YeeArgs<double>> &, hipStream_t);
template void launch_h_batch<float>(const BatchLaunch<YeeArgs<float>> &, hipStream_t);
template void launch_de_batch<double>(const BatchLaunch<YeeArgs<double>> &, int, hipStream_t);
template void launch_de_batch<float>(const BatchLaunch<YeeArgs<float>> &, int, hipStream_t);

}  // namespace phpfdtd
