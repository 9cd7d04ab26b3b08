// kernels_yee.hip - H update and fused D/E update for gfx950.
//
// Work decomposition: a wavefront owns a 64-cell x segment of R consecutive y
// rows and marches up a chunk of z planes.  Per plane it loads each field row
// once: the y neighbour of row r is row r+1 (H) or r-1 (D) held in registers,
// so only one halo row per R rows is fetched twice; the z neighbour (k+1 of E
// for H, k-1 of H for D) is carried from the previous plane.  The x neighbour
// (i+1 / i-1) is a load of the same row shifted by one cell, served by the L1
// from the lines the row load just brought in.  j and k are wavefront-uniform,
// so every field access is one coalesced row segment.
//
// Hardware places work-group b on XCD (b mod 8).  xcd_tile() renumbers the
// groups so that each XCD owns one contiguous range of tiles; neighbouring
// tiles then share the XCD's L2.
//
// Reference semantics (fdtd.c):
//   H: i < nx-1, j < ny-1, all k; E above the top plane is 0 (3123-3231).
//   D: 1 <= i <= nx-2, 1 <= j <= ny-2, all k; H below k=0 is 0 (3238-3356).
//   CPML psi is lagged: the old psi enters the field, then psi is advanced
//   (Q1).  psi is stored pre-multiplied by dt.
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

template <typename T>
__device__ __forceinline__ T ldnt(const T *p) {
    return __builtin_nontemporal_load(p);
}

template <typename T>
__device__ __forceinline__ void stnt(T v, T *p) {
    __builtin_nontemporal_store(v, p);
}

template <typename real, int R>
__global__ __launch_bounds__(256) void h_update(YeeArgs<real> a) {
    const int t = xcd_tile(blockIdx.x, a.nblocks, a.nblocks_pad);
    if (t < 0) return;
    const int bx = t % a.nbx;
    const int rest = t / a.nbx;
    const int by = rest % a.nby;
    const int bz = rest / a.nby;
    const int i = bx * 64 + threadIdx.x;
    const int j0 = uniform((by * kBlockY + threadIdx.y) * R);
    const int nr = min(R, a.ny - 1 - j0);  // H rows j < ny-1 of this wavefront
    if (nr <= 0 || i >= a.nx - 1) return;
    const int k0 = uniform(bz * a.kchunk);
    const int k1 = min(k0 + a.kchunk, a.nz);
    const size_t nxp = a.nxp;
    const size_t plane = nxp * a.ny;
    const int L2 = 2 * a.L;

    const real khx = a.ax.kh[i];
    const int px = a.ax.slot[i];
    real bxv = 0, chx = 0;
    if (px >= 0) { bxv = a.ax.b[i]; chx = a.ax.ch[i]; }
    real khy[R];
    int py[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int j = min(j0 + r, a.ny - 1);
        khy[r] = a.ay.kh[j];
        py[r] = uniform(a.ay.slot[j]);
    }

    // plane k rows j0..j0+nr of Ex (row nr is the y halo), rows j0..j0+nr-1 of Ey
    size_t c0 = ((size_t)k0 * a.ny + j0) * nxp + i;
    real ex[R + 1], ey[R];
#pragma unroll
    for (int r = 0; r <= R; ++r)
        if (r <= nr) ex[r] = a.Ex[c0 + r * nxp];
#pragma unroll
    for (int r = 0; r < R; ++r)
        if (r < nr) ey[r] = a.Ey[c0 + r * nxp];

    for (int k = k0; k < k1; ++k, c0 += plane) {
        real exn[R + 1], eyn[R], ez[R + 1];
        const bool top = k + 1 >= a.nz;  // E above the top plane is the zero halo
#pragma unroll
        for (int r = 0; r <= R; ++r)
            if (r <= nr) {
                exn[r] = top ? real(0) : a.Ex[c0 + plane + r * nxp];
                ez[r] = a.Ez[c0 + r * nxp];
            }
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (r < nr) eyn[r] = top ? real(0) : a.Ey[c0 + plane + r * nxp];
        const real khz = a.az.kh[k];
        const int pz = uniform(a.az.slot[k]);
        real bzv = 0, chz = 0;
        if (pz >= 0) { bzv = a.az.b[k]; chz = a.az.ch[k]; }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (r >= nr) continue;
            const size_t c = c0 + r * nxp;
            const int j = j0 + r;
            const real ez_ip = a.Ez[c + 1], ey_ip = a.Ey[c + 1];
            const real dyEz = ez[r + 1] - ez[r], dzEy = eyn[r] - ey[r];
            const real dxEz = ez_ip - ez[r], dzEx = exn[r] - ex[r];
            const real dxEy = ey_ip - ey[r], dyEx = ex[r + 1] - ex[r];
            real hx = ldnt(a.Hx + c), hy = ldnt(a.Hy + c), hz = ldnt(a.Hz + c);
            hx += khz * dzEy - khy[r] * dyEz;
            hy += khx * dxEz - khz * dzEx;
            hz += khy[r] * dyEx - khx * dxEy;
            if (px >= 0) {
                const size_t q = ((size_t)k * a.ny + j) * L2 + px;
                const real sxy = a.pHxy[q], sxz = a.pHxz[q];
                hy += sxy;
                hz += sxz;
                a.pHxy[q] = bxv * sxy + chx * dxEz;
                a.pHxz[q] = bxv * sxz - chx * dxEy;
            }
            if (py[r] >= 0) {
                const size_t q = ((size_t)k * L2 + py[r]) * nxp + i;
                const real byv = a.ay.b[j], chy = a.ay.ch[j];
                const real syx = a.pHyx[q], syz = a.pHyz[q];
                hx += syx;
                hz += syz;
                a.pHyx[q] = byv * syx - chy * dyEz;
                a.pHyz[q] = byv * syz + chy * dyEx;
            }
            if (pz >= 0) {
                const size_t q = ((size_t)pz * a.ny + j) * nxp + i;
                const real szx = a.pHzx[q], szy = a.pHzy[q];
                hx += szx;
                hy += szy;
                a.pHzx[q] = bzv * szx + chz * dzEy;
                a.pHzy[q] = bzv * szy - chz * dzEx;
            }
            stnt(hx, a.Hx + c);
            stnt(hy, a.Hy + c);
            stnt(hz, a.Hz + c);
        }
#pragma unroll
        for (int r = 0; r <= R; ++r) ex[r] = exn[r];
#pragma unroll
        for (int r = 0; r < R; ++r) ey[r] = eyn[r];
    }
}

template <typename real, int R>
__global__ __launch_bounds__(256) void de_update(YeeArgs<real> a, int nmat) {
    extern __shared__ unsigned char lds_raw[];
    MatCoef<real> *mats = reinterpret_cast<MatCoef<real> *>(lds_raw);
    {
        const int nw = nmat * (int)(sizeof(MatCoef<real>) / 4);
        const uint32_t *src = reinterpret_cast<const uint32_t *>(a.mats);
        uint32_t *dst = reinterpret_cast<uint32_t *>(lds_raw);
        for (int w = threadIdx.y * 64 + threadIdx.x; w < nw; w += 256) dst[w] = src[w];
        __syncthreads();
    }
    const int t = xcd_tile(blockIdx.x, a.nblocks, a.nblocks_pad);
    if (t < 0) return;
    const int bx = t % a.nbx;
    const int rest = t / a.nbx;
    const int by = rest % a.nby;
    const int bz = rest / a.nby;
    const int i = bx * 64 + threadIdx.x;
    const int j0 = uniform((by * kBlockY + threadIdx.y) * R);
    const int nr = min(R, a.ny - j0);  // E/D rows of this wavefront
    if (nr <= 0 || i >= a.nx) return;
    const int k0 = uniform(bz * a.kchunk);
    const int k1 = min(k0 + a.kchunk, a.nz);
    const size_t nxp = a.nxp;
    const size_t plane = nxp * a.ny;
    const int L2 = 2 * a.L;
    const bool inx = i >= 1 && i <= a.nx - 2;
    // rows of the D loop: 1 <= j <= ny-2 (wavefront-uniform per row)
    bool iny[R];
    real kdy[R];
    int py[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int j = j0 + r;
        iny[r] = j >= 1 && j <= a.ny - 2 && r < nr;
        const int jc = min(j, a.ny - 1);
        kdy[r] = a.ay.kd[jc];
        py[r] = uniform(a.ay.slot[jc]);
    }
    const bool halo = j0 >= 1;  // row j0-1 exists
    const real kdx = a.ax.kd[i];
    const int px = a.ax.slot[i];
    real bxv = 0, cdx = 0;
    if (px >= 0) { bxv = a.ax.b[i]; cdx = a.ax.cd[i]; }

    // step index selects this step's source values (value*dt), fdtd.c:3363-3396
    const int64_t st = a.nsrc > 0 ? *a.step : 0;

    size_t c0 = ((size_t)k0 * a.ny + j0) * nxp + i;
    // H at plane k-1 (rows of this wavefront), zero below k = 0
    real hxm[R], hym[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        hxm[r] = 0;
        hym[r] = 0;
        if (r < nr && k0 > 0) {
            hxm[r] = a.Hx[c0 - plane + r * nxp];
            hym[r] = a.Hy[c0 - plane + r * nxp];
        }
    }
    for (int k = k0; k < k1; ++k, c0 += plane) {
        // rows -1..nr-1 of Hx and Hz (row -1 is the y halo), rows 0..nr-1 of Hy
        real hx[R + 1], hz[R + 1], hy[R];
        hx[0] = 0;
        hz[0] = 0;
        if (halo) {
            hx[0] = a.Hx[c0 - nxp];
            hz[0] = a.Hz[c0 - nxp];
        }
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (r < nr) {
                hx[r + 1] = a.Hx[c0 + r * nxp];
                hz[r + 1] = a.Hz[c0 + r * nxp];
                hy[r] = a.Hy[c0 + r * nxp];
            }
        const real kdz = a.az.kd[k];
        const int pz = uniform(a.az.slot[k]);
        real bzv = 0, cdz = 0;
        if (pz >= 0) { bzv = a.az.b[k]; cdz = a.az.cd[k]; }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (r >= nr) continue;
            const size_t c = c0 + r * nxp;
            const int j = j0 + r;
            real dx = ldnt(a.Dx + c), dy = ldnt(a.Dy + c), dz = ldnt(a.Dz + c);
            const int m = a.medium[c];
            const int fl = mats[m].flags;
            if (inx && iny[r]) {
                const real hy_im = a.Hy[c - 1], hz_im = a.Hz[c - 1];
                const real dyHz = hz[r + 1] - hz[r], dzHy = hy[r] - hym[r];
                const real dzHx = hx[r + 1] - hxm[r], dxHz = hz[r + 1] - hz_im;
                const real dxHy = hy[r] - hy_im, dyHx = hx[r + 1] - hx[r];
                dx += kdy[r] * dyHz - kdz * dzHy;
                dy += kdz * dzHx - kdx * dxHz;
                dz += kdx * dxHy - kdy[r] * dyHx;
                if (px >= 0) {
                    const size_t q = ((size_t)k * a.ny + j) * L2 + px;
                    const real sxy = a.pDxy[q], sxz = a.pDxz[q];
                    dy += sxy;
                    dz += sxz;
                    a.pDxy[q] = bxv * sxy - cdx * dxHz;
                    a.pDxz[q] = bxv * sxz + cdx * dxHy;
                }
                if (py[r] >= 0) {
                    const size_t q = ((size_t)k * L2 + py[r]) * nxp + i;
                    const real byv = a.ay.b[j], cdy = a.ay.cd[j];
                    const real syx = a.pDyx[q], syz = a.pDyz[q];
                    dx += syx;
                    dz += syz;
                    a.pDyx[q] = byv * syx + cdy * dyHz;
                    a.pDyz[q] = byv * syz - cdy * dyHx;
                }
                if (pz >= 0) {
                    const size_t q = ((size_t)pz * a.ny + j) * nxp + i;
                    const real szx = a.pDzx[q], szy = a.pDzy[q];
                    dx += szx;
                    dy += szy;
                    a.pDzx[q] = bzv * szx - cdz * dzHy;
                    a.pDzy[q] = bzv * szy + cdz * dzHx;
                }
                if (fl & MF_LOSS) {
                    const real ex = a.Ex[c], ey = a.Ey[c], ez = a.Ez[c];
                    const real *Lk = mats[m].Lk;
                    dx -= Lk[0] * ex + Lk[1] * ey + Lk[2] * ez;
                    dy -= Lk[3] * ex + Lk[4] * ey + Lk[5] * ez;
                    dz -= Lk[6] * ex + Lk[7] * ey + Lk[8] * ez;
                }
            }
            for (int s = 0; s < a.nsrc; ++s) {
                if (a.src_k[s] == k && a.src_j[s] == j && a.src_i[s] == i) {
                    const real v = a.src_val[st * a.nsrc + s];
                    if (a.src_c[s] == 0) dx += v;
                    else if (a.src_c[s] == 1) dy += v;
                    else dz += v;
                }
            }
            stnt(dx, a.Dx + c);
            stnt(dy, a.Dy + c);
            stnt(dz, a.Dz + c);
            if (!(fl & MF_SPECIAL)) {
                const real *M = mats[m].M;
                stnt(M[0] * dx + M[1] * dy + M[2] * dz, a.Ex + c);
                stnt(M[3] * dx + M[4] * dy + M[5] * dz, a.Ey + c);
                stnt(M[6] * dx + M[7] * dy + M[8] * dz, a.Ez + c);
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            hxm[r] = hx[r + 1];
            hym[r] = hy[r];
        }
    }
}

template <typename real>
void launch_h(const YeeArgs<real> &a, hipStream_t s) {
    hipLaunchKernelGGL((h_update<real, YeeRows<real>::value>), dim3(a.nblocks_pad), dim3(64, kBlockY), 0, s, a);
}

template <typename real>
void launch_de(const YeeArgs<real> &a, int nmat, hipStream_t s) {
    const size_t lds = (size_t)nmat * sizeof(MatCoef<real>);
    hipLaunchKernelGGL((de_update<real, YeeRows<real>::value>), dim3(a.nblocks_pad), dim3(64, kBlockY), lds, s, a,
                       nmat);
}

template void launch_h<double>(const YeeArgs<double> &, hipStream_t);
template void launch_h<float>(const YeeArgs<float> &, hipStream_t);
template void launch_de<double>(const YeeArgs<double> &, int, hipStream_t);
template void launch_de<float>(const YeeArgs<float> &, int, hipStream_t);

}  // namespace phpfdtd
