// kernels_yee.hip - H update and fused D/E update for gfx950.
//
// Work decomposition: a thread owns one (i, j) column and marches up a chunk
// of z planes, so the z neighbour (k+1 of E for H, k-1 of H for D) is carried
// in registers instead of being re-read.  Work-groups are 64 x 4 threads: a
// wavefront is one 64-cell x row segment of one j row, so j and k are
// wavefront-uniform and every field access is one coalesced row segment.  The
// x and y neighbours come from the L1/L2 copies of the rows the neighbouring
// lanes and wavefronts load in the same iteration.
//
// Hardware places work-group b on XCD (b mod 8).  xcd_tile() renumbers the
// groups so that each XCD owns one contiguous range of tiles; y-neighbour
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

template <typename real>
__global__ __launch_bounds__(256) void h_update(YeeArgs<real> a) {
    const int t = xcd_tile(blockIdx.x, a.nblocks, a.nblocks_pad);
    if (t < 0) return;
    const int bx = t % a.nbx;
    const int rest = t / a.nbx;
    const int by = rest % a.nby;
    const int bz = rest / a.nby;
    const int i = bx * 64 + threadIdx.x;
    const int j = uniform(by * kBlockY + threadIdx.y);
    if (j >= a.ny - 1 || i >= a.nx - 1) return;
    const int k0 = uniform(bz * a.kchunk);
    const int k1 = min(k0 + a.kchunk, a.nz);
    const size_t nxp = a.nxp;
    const size_t plane = nxp * a.ny;
    const int L2 = 2 * a.L;
    size_t c = ((size_t)k0 * a.ny + j) * nxp + i;

    const real khx = a.ax.kh[i], khy = a.ay.kh[j];
    const int px = a.ax.slot[i], py = uniform(a.ay.slot[j]);
    real bxv = 0, chx = 0, byv = 0, chy = 0;
    if (px >= 0) { bxv = a.ax.b[i]; chx = a.ax.ch[i]; }
    if (py >= 0) { byv = a.ay.b[j]; chy = a.ay.ch[j]; }

    real ex_k = a.Ex[c], ey_k = a.Ey[c];
    for (int k = k0; k < k1; ++k, c += plane) {
        real ex_kp = 0, ey_kp = 0;
        if (k + 1 < a.nz) { ex_kp = a.Ex[c + plane]; ey_kp = a.Ey[c + plane]; }
        const real ez = a.Ez[c];
        const real ez_ip = a.Ez[c + 1], ey_ip = a.Ey[c + 1];
        const real ez_jp = a.Ez[c + nxp], ex_jp = a.Ex[c + nxp];
        const real dyEz = ez_jp - ez, dzEy = ey_kp - ey_k;
        const real dxEz = ez_ip - ez, dzEx = ex_kp - ex_k;
        const real dxEy = ey_ip - ey_k, dyEx = ex_jp - ex_k;
        const real khz = a.az.kh[k];
        real hx = a.Hx[c], hy = a.Hy[c], hz = a.Hz[c];
        hx += khz * dzEy - khy * dyEz;
        hy += khx * dxEz - khz * dzEx;
        hz += khy * dyEx - khx * dxEy;
        if (px >= 0) {
            const size_t q = ((size_t)k * a.ny + j) * L2 + px;
            const real sxy = a.pHxy[q], sxz = a.pHxz[q];
            hy += sxy;
            hz += sxz;
            a.pHxy[q] = bxv * sxy + chx * dxEz;
            a.pHxz[q] = bxv * sxz - chx * dxEy;
        }
        if (py >= 0) {
            const size_t q = ((size_t)k * L2 + py) * nxp + i;
            const real syx = a.pHyx[q], syz = a.pHyz[q];
            hx += syx;
            hz += syz;
            a.pHyx[q] = byv * syx - chy * dyEz;
            a.pHyz[q] = byv * syz + chy * dyEx;
        }
        const int pz = uniform(a.az.slot[k]);
        if (pz >= 0) {
            const size_t q = ((size_t)pz * a.ny + j) * nxp + i;
            const real bzv = a.az.b[k], chz = a.az.ch[k];
            const real szx = a.pHzx[q], szy = a.pHzy[q];
            hx += szx;
            hy += szy;
            a.pHzx[q] = bzv * szx + chz * dzEy;
            a.pHzy[q] = bzv * szy - chz * dzEx;
        }
        a.Hx[c] = hx;
        a.Hy[c] = hy;
        a.Hz[c] = hz;
        ex_k = ex_kp;
        ey_k = ey_kp;
    }
}

template <typename real>
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
    const int j = uniform(by * kBlockY + threadIdx.y);
    if (j >= a.ny || i >= a.nx) return;
    const int k0 = uniform(bz * a.kchunk);
    const int k1 = min(k0 + a.kchunk, a.nz);
    const size_t nxp = a.nxp;
    const size_t plane = nxp * a.ny;
    const int L2 = 2 * a.L;
    const bool interior = (i >= 1 && i <= a.nx - 2 && j >= 1 && j <= a.ny - 2);
    size_t c = ((size_t)k0 * a.ny + j) * nxp + i;

    const real kdx = a.ax.kd[i], kdy = a.ay.kd[j];
    const int px = a.ax.slot[i], py = uniform(a.ay.slot[j]);
    real bxv = 0, cdx = 0, byv = 0, cdy = 0;
    if (px >= 0) { bxv = a.ax.b[i]; cdx = a.ax.cd[i]; }
    if (py >= 0) { byv = a.ay.b[j]; cdy = a.ay.cd[j]; }

    // step index selects this step's source values (value*dt), fdtd.c:3363-3396
    const int64_t st = a.nsrc > 0 ? *a.step : 0;

    real hx_km = 0, hy_km = 0;
    if (interior && k0 > 0) { hx_km = a.Hx[c - plane]; hy_km = a.Hy[c - plane]; }
    for (int k = k0; k < k1; ++k, c += plane) {
        real dx = a.Dx[c], dy = a.Dy[c], dz = a.Dz[c];
        const int m = a.medium[c];
        const int fl = mats[m].flags;
        if (interior) {
            const real hx = a.Hx[c], hy = a.Hy[c], hz = a.Hz[c];
            const real hy_im = a.Hy[c - 1], hz_im = a.Hz[c - 1];
            const real hx_jm = a.Hx[c - nxp], hz_jm = a.Hz[c - nxp];
            const real dyHz = hz - hz_jm, dzHy = hy - hy_km;
            const real dzHx = hx - hx_km, dxHz = hz - hz_im;
            const real dxHy = hy - hy_im, dyHx = hx - hx_jm;
            const real kdz = a.az.kd[k];
            dx += kdy * dyHz - kdz * dzHy;
            dy += kdz * dzHx - kdx * dxHz;
            dz += kdx * dxHy - kdy * dyHx;
            if (px >= 0) {
                const size_t q = ((size_t)k * a.ny + j) * L2 + px;
                const real sxy = a.pDxy[q], sxz = a.pDxz[q];
                dy += sxy;
                dz += sxz;
                a.pDxy[q] = bxv * sxy - cdx * dxHz;
                a.pDxz[q] = bxv * sxz + cdx * dxHy;
            }
            if (py >= 0) {
                const size_t q = ((size_t)k * L2 + py) * nxp + i;
                const real syx = a.pDyx[q], syz = a.pDyz[q];
                dx += syx;
                dz += syz;
                a.pDyx[q] = byv * syx + cdy * dyHz;
                a.pDyz[q] = byv * syz - cdy * dyHx;
            }
            const int pz = uniform(a.az.slot[k]);
            if (pz >= 0) {
                const size_t q = ((size_t)pz * a.ny + j) * nxp + i;
                const real bzv = a.az.b[k], cdz = a.az.cd[k];
                const real szx = a.pDzx[q], szy = a.pDzy[q];
                dx += szx;
                dy += szy;
                a.pDzx[q] = bzv * szx - cdz * dzHy;
                a.pDzy[q] = bzv * szy + cdz * dzHx;
            }
            hx_km = hx;
            hy_km = hy;
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
        a.Dx[c] = dx;
        a.Dy[c] = dy;
        a.Dz[c] = dz;
        if (!(fl & MF_SPECIAL)) {
            const real *M = mats[m].M;
            a.Ex[c] = M[0] * dx + M[1] * dy + M[2] * dz;
            a.Ey[c] = M[3] * dx + M[4] * dy + M[5] * dz;
            a.Ez[c] = M[6] * dx + M[7] * dy + M[8] * dz;
        }
    }
}

template <typename real>
void launch_h(const YeeArgs<real> &a, hipStream_t s) {
    hipLaunchKernelGGL(h_update<real>, dim3(a.nblocks_pad), dim3(64, kBlockY), 0, s, a);
}

template <typename real>
void launch_de(const YeeArgs<real> &a, int nmat, hipStream_t s) {
    const size_t lds = (size_t)nmat * sizeof(MatCoef<real>);
    hipLaunchKernelGGL(de_update<real>, dim3(a.nblocks_pad), dim3(64, kBlockY), lds, s, a, nmat);
}

template void launch_h<double>(const YeeArgs<double> &, hipStream_t);
template void launch_h<float>(const YeeArgs<float> &, hipStream_t);
template void launch_de<double>(const YeeArgs<double> &, int, hipStream_t);
template void launch_de<float>(const YeeArgs<float> &, int, hipStream_t);

}  // namespace phpfdtd
