// kernels_aux.hip - per-step epilogue, diagnostics and sampling kernels.
#include <algorithm>

#include "phpfdtd_internal.h"

namespace phpfdtd {

// Per-step epilogue (one thread): the monitor row of fdtd.c:3806-3851 for the
// step that just finished, then the device step counter advances.  The row is
// written at index `step`, so graph replays need no host bookkeeping; step
// s wrote E/H set (s + 1) & 1, which the row samples.  part 0 writes the
// whole row; part 1 only D and H, part 2 t, E, R, V and advances the step.
__device__ __forceinline__ void step_row(const StepEndArgs &a, bool dh, bool rest, bool f64) {
    const int64_t st = *a.step;
    const void *const *F = a.F[(st + 1) & 1];
    if (a.mon_on && st < a.ring_rows) {
        double *row = a.ring + st * 12;
        auto val = [&](int q) -> double {
            if (a.mon_c < 0) return 0.0;
            return f64 ? static_cast<const double *>(F[q])[a.mon_c] : (double)static_cast<const float *>(F[q])[a.mon_c];
        };
        if (rest) {
            row[0] = (double)st * a.dt;
            for (int q = 0; q < 3; ++q) row[1 + q] = val(q);
            double rv[2] = {0.0, 0.0};
            if (a.mon_ridx >= 0) {
                if (a.mon_rc >= 0) {  // compensated FP32 state
                    rv[0] = pk_unpack(static_cast<const float *>(a.mon_r)[a.mon_ridx], pk_code_at(a.mon_r, a.mon_rc));
                    rv[1] = pk_unpack(static_cast<const float *>(a.mon_v)[a.mon_vidx], pk_code_at(a.mon_vcb, a.mon_vc));
                } else {
                    rv[0] = static_cast<const double *>(a.mon_r)[a.mon_ridx];
                    rv[1] = static_cast<const double *>(a.mon_v)[a.mon_vidx];
                }
            }
            row[10] = rv[0];
            row[11] = rv[1];
        }
        if (dh) {
            for (int q = 3; q < 9; ++q) row[1 + q] = val(q);
            if (a.mon_dfe && a.mon_c >= 0) {  // E-form plain cell: D = M^-1 E, as derive_d computes it
                const double e0 = val(0), e1 = val(1), e2 = val(2);
                for (int r = 0; r < 3; ++r) {
                    const double *m = a.mon_dm + 3 * r;
                    const double d = fma(m[2], e2, fma(m[1], e1, m[0] * e0));
                    row[4 + r] = f64 ? d : (double)(float)d;
                }
            }
        }
    }
    if (rest) *a.step = st + 1;
}

template <typename real>
__global__ void step_end(StepEndArgs a) {
    if (threadIdx.x != 0) return;
    step_row(a, a.part != 2, a.part != 1, sizeof(real) == 8);
}

// step_end of every domain of a batch: work-group b handles domain b.
template <typename real>
__global__ void step_end_batch(const StepEndArgs *args, int n) {
    if (threadIdx.x != 0 || (int)blockIdx.x >= n) return;
    const StepEndArgs a = args[blockIdx.x];
    step_row(a, true, true, sizeof(real) == 8);
}

// D = M^-1 . E for the E-form plain cells of a box; their D is not stored by
// the time step (D of other cells is, and is left as it is).
template <typename real>
__global__ __launch_bounds__(256) void derive_d(DeriveArgs a) {
    const int set = (int)((*a.step + a.ahead) & 1);
    const real *Ex = static_cast<const real *>(a.E[set][0]), *Ey = static_cast<const real *>(a.E[set][1]),
               *Ez = static_cast<const real *>(a.E[set][2]);
    const int64_t bx = a.i1 - a.i0, by = a.j1 - a.j0, bz = a.k1 - a.k0;
    const int64_t n = bx * by * bz;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = a.i0 + q % bx, j = a.j0 + (q / bx) % by, k = a.k0 + q / (bx * by);
        const size_t c = ((size_t)k * a.ny + j) * a.nxp + i;
        const int m = a.medium[c];
        if (!a.derive[m]) continue;
        const double e0 = Ex[c], e1 = Ey[c], e2 = Ez[c];
        for (int r = 0; r < 3; ++r) {
            const double *dm = a.dm + 9 * m + 3 * r;
            static_cast<real *>(a.D[r])[c] = (real)fma(dm[2], e2, fma(dm[1], e1, dm[0] * e0));
        }
    }
}

template <typename T>
__global__ void fill(T *p, T v, int64_t n) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        p[q] = v;
}

__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(v, off, 64);
        v = v > o ? v : o;
    }
    return v;
}

// max |E| and max |Ex| over the physical cells (fdtd.c:3537-3552).  Values are
// non-negative, so their IEEE bit patterns order like the values.  A
// work-group takes whole rows of the padded arrays (the row index is
// wavefront-uniform), so no per-element division is needed.
template <typename real>
__global__ __launch_bounds__(256) void reduce_e(const real *Ex, const real *Ey, const real *Ez, int nx, int ny,
                                                int nz, int nxp, unsigned long long *out) {
    const int64_t rows = (int64_t)ny * nz;
    unsigned long long me = 0, mx = 0;
    for (int64_t row = blockIdx.x; row < rows; row += gridDim.x) {
        const int64_t base = row * nxp;
        for (int i = threadIdx.x; i < nx; i += blockDim.x) {
            const double ex = Ex[base + i], ey = Ey[base + i], ez = Ez[base + i];
            const double e = sqrt(ex * ex + ey * ey + ez * ez);
            const unsigned long long be = __double_as_longlong(e), bx = __double_as_longlong(fabs(ex));
            me = me > be ? me : be;
            mx = mx > bx ? mx : bx;
        }
    }
    me = wave_max_u64(me);
    mx = wave_max_u64(mx);
    if ((threadIdx.x & 63) == 0) {
        atomicMax(out + 0, me);
        atomicMax(out + 1, mx);
    }
}

// (max |r| bits, smallest key among the dofs holding it): the larger value
// wins, equal values keep the smaller key.
__device__ __forceinline__ void pair_merge(unsigned long long &b, unsigned long long &k, unsigned long long ob,
                                           unsigned long long ok) {
    if (ob > b || (ob == b && ok < k)) {
        b = ob;
        k = ok;
    }
}

// max |r| of one lattice material and its first argmax in the reference's
// loop order (i, j, k, atom, dir) (fdtd.c:3569-3592), in one pass over r.
// A wavefront takes a 64-slot tile as the lattice kernel does: lane = slot,
// one 16-byte vector per dof group, so every load is one 1 KiB row.  Values
// are the stored values themselves: FP64, or compensated FP32 unpacked
// exactly (word + byte).  The key of dof q of cell (i, j, k) is
// ((i ny + j) nz + k) 96 + q, which orders like the reference's loops; as
// there, a value must exceed the running maximum strictly (r_val >
// local_max_r, starting from 0), so zeros and NaNs never become the argmax.
// Each wavefront leaves its pair in part[2 w], part[2 w + 1].
template <bool PK>
__global__ __launch_bounds__(256) void diag_r(const void *r, const uint32_t *cells, int64_t nslot, IonLayout lay,
                                              int ny, int nz, int nxp, unsigned long long *part) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwave = ((int64_t)gridDim.x * blockDim.x) >> 6;
    const int64_t ntile = (nslot + 63) / 64;
    const int ng = lay.ng(), n3 = lay.n3;
    unsigned long long best = 0, key = ~0ull;
    // a lane visits its slots in device (x-fastest) order, which is not the
    // reference's (i, j, k) order: values merge by key, and only positive
    // values count (the reference's strict r_val > local_max_r from 0)
    auto take = [&](double v, unsigned long long k) {
        const unsigned long long bits = __double_as_longlong(fabs(v));
        if (v == v && bits > 0) pair_merge(best, key, bits, k);
    };
    for (int64_t t = wave; t < ntile; t += nwave) {
        const int64_t s = t * 64 + lane;
        if (s >= nslot) continue;
        const int64_t c = cells[s];
        const int64_t ci = c % nxp, cj = (c / nxp) % ny, ck = c / ((int64_t)nxp * ny);
        const unsigned long long base = (unsigned long long)(((ci * ny + cj) * nz + ck) * 96);
        if (PK) {
            // group g: words w[g][lane][0..3]; its 4-bit codes are bits
            // 16 (g % 2) .. +15 of dword (g % 8) / 2 of lane's 16 bytes in
            // compensation row g / 8
            const float *tile = static_cast<const float *>(r) + t * lay.tile_r();
            const float4 *w = reinterpret_cast<const float4 *>(tile);
            const int4 *cb = reinterpret_cast<const int4 *>(tile + (int64_t)kPkWords * ng);
            for (int g0 = 0; g0 < ng; g0 += 8) {
                const int4 bv = cb[(g0 >> 3) * 64 + lane];
                const uint32_t bws[4] = {(uint32_t)bv.x, (uint32_t)bv.y, (uint32_t)bv.z, (uint32_t)bv.w};
#pragma unroll
                for (int h = 0; h < 8; ++h) {
                    const int g = g0 + h;
                    if (g >= ng) break;
                    const float4 hv = w[g * 64 + lane];
                    const float hw[4] = {hv.x, hv.y, hv.z, hv.w};
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        if (4 * g + e < n3)
                            take(pk_unpack(hw[e], pk_code_of(bws[h / 2], 16 * (h % 2) + 4 * e)), base + 4 * g + e);
                }
            }
        } else {
            const double2 *w = reinterpret_cast<const double2 *>(static_cast<const double *>(r) + t * lay.tile_r());
            for (int g = 0; g < ng; ++g) {
                const double2 hv = w[g * 64 + lane];
                if (2 * g < n3) take(hv.x, base + 2 * g);
                if (2 * g + 1 < n3) take(hv.y, base + 2 * g + 1);
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long ob = __shfl_xor(best, off, 64), ok = __shfl_xor(key, off, 64);
        pair_merge(best, key, ob, ok);
    }
    if (lane == 0) {
        part[2 * wave] = best;
        part[2 * wave + 1] = key;
    }
}

// Merge n wave pairs into out[2] (max |r| bits) and out[3] (its key); one
// work-group, run after each diag_r on the same stream.
__global__ __launch_bounds__(256) void diag_r_merge(const unsigned long long *part, int64_t n, unsigned long long *out) {
    __shared__ unsigned long long sb[4], sk[4];
    unsigned long long best = 0, key = ~0ull;
    for (int64_t w = threadIdx.x; w < n; w += blockDim.x) pair_merge(best, key, part[2 * w], part[2 * w + 1]);
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long ob = __shfl_xor(best, off, 64), ok = __shfl_xor(key, off, 64);
        pair_merge(best, key, ob, ok);
    }
    if ((threadIdx.x & 63) == 0) {
        sb[threadIdx.x >> 6] = best;
        sk[threadIdx.x >> 6] = key;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long b = out[2], k = out[3];
        for (int q = 0; q < 4; ++q) pair_merge(b, k, sb[q], sk[q]);
        out[2] = b;
        out[3] = k;
    }
}

// Field index of element q of a plane in the reference's slice layout
// (fdtd.c:999-1400): axis 'z' -> [nx][ny], 'x' -> [ny][nz], 'y' -> [nx][nz].
// Returns -1 past the end of the plane.
__device__ __forceinline__ int64_t plane_cell(int64_t q, char axis, int pos, int nx, int ny, int nz, int nxp) {
    if (axis == 'z') {
        if (q >= (int64_t)nx * ny) return -1;
        const int i = (int)(q / ny), j = (int)(q % ny);
        return ((int64_t)pos * ny + j) * nxp + i;
    }
    if (axis == 'x') {
        if (q >= (int64_t)ny * nz) return -1;
        const int j = (int)(q / nz), k = (int)(q % nz);
        return ((int64_t)k * ny + j) * nxp + pos;
    }
    if (q >= (int64_t)nx * nz) return -1;
    const int i = (int)(q / nz), k = (int)(q % nz);
    return ((int64_t)k * ny + pos) * nxp + i;
}

// Value of a component at field index c in FP64; emag selects |E| computed in
// FP64 from Ex, Ey, Ez (fdtd.c:3693-3703).
template <typename real>
__device__ __forceinline__ double sample_value(const real *F, const real *Ex, const real *Ey, const real *Ez, int emag,
                                               int64_t c) {
    if (emag) {
        const double ex = Ex[c], ey = Ey[c], ez = Ez[c];
        return sqrt(ex * ex + ey * ey + ez * ez);
    }
    return (double)F[c];
}

// One plane as float32 in the reference's slice layout.
template <typename real>
__global__ __launch_bounds__(256) void slice_k(const real *F, const real *Ex, const real *Ey, const real *Ez,
                                               int emag, char axis, int pos, int nx, int ny, int nz, int nxp,
                                               float *out) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t c = plane_cell(q, axis, pos, nx, ny, nz, nxp);
    if (c < 0) return;
    out[q] = (float)sample_value(F, Ex, Ey, Ez, emag, c);
}

// Streamed single-bin DFT of one plane (frequency_analysis.c:548-569 for one
// bin): at steps t0 + n*interval, n < nsamp, acc[q] += x_n(q) * w[n] with
// w[n] = exp(-2 pi i k n / N) / N precomputed in FP64.  ref_float rounds the
// sample to the float32 a slice file holds.  Runs inside the step graph
// before the step counter advances, so it sees the step just updated (in
// E/H set (step + 1) & 1).
template <typename real>
__global__ __launch_bounds__(256) void dft_accum(DftArgs a) {
    const int64_t st = *a.step - a.t0;
    if (st < 0 || st % a.interval) return;
    const int64_t n = st / a.interval;
    if (n >= a.nsamp) return;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t c = plane_cell(q, a.axis, a.pos, a.nx, a.ny, a.nz, a.nxp);
    if (c < 0) return;
    const int set = (int)((st + a.t0 + 1) & 1);
    const real *F = static_cast<const real *>(a.F[set]);
    const real *Ex = static_cast<const real *>(a.Ex[set]), *Ey = static_cast<const real *>(a.Ey[set]),
               *Ez = static_cast<const real *>(a.Ez[set]);
    double x = c < a.ncell ? sample_value(F, Ex, Ey, Ez, a.emag, c) : 0.0;
    if (a.ref_float) x = (double)(float)x;
    const double2 w = a.w[n];
    double2 acc = a.acc[q];
    acc.x = fma(x, w.x, acc.x);
    acc.y = fma(x, w.y, acc.y);
    a.acc[q] = acc;
}

template <typename real>
void launch_step_end(const StepEndArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(step_end<real>, dim3(1), dim3(64), 0, s, a);
}

template <typename real>
void launch_step_end_batch(const StepEndArgs *args, int n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(step_end_batch<real>, dim3(n), dim3(64), 0, s, args, n);
}

template <typename T>
void launch_fill(T *p, T v, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    const int64_t nb = (n + 255) / 256;
    hipLaunchKernelGGL(fill<T>, dim3((unsigned)(nb < 4096 ? nb : 4096)), dim3(256), 0, s, p, v, n);
}

template <typename real>
void launch_reduce_e(const real *Ex, const real *Ey, const real *Ez, int nx, int ny, int nz, int nxp,
                     unsigned long long *out, hipStream_t s) {
    const int64_t rows = (int64_t)ny * nz;
    hipLaunchKernelGGL(reduce_e<real>, dim3((unsigned)std::min<int64_t>(rows, 4096)), dim3(256), 0, s, Ex, Ey, Ez,
                       nx, ny, nz, nxp, out);
}

// max |r| and its first argmax over one lattice material (r: layout lay,
// compensated FP32 when pk), merged into out[2], out[3]; part holds
// kDiagWaves pairs.
void launch_diag_r(const void *r, bool pk, const uint32_t *cells, int64_t nslot, IonLayout lay, int ny, int nz,
                   int nxp, unsigned long long *part, unsigned long long *out, hipStream_t s) {
    if (nslot <= 0) return;
    const int64_t ntile = (nslot + 63) / 64;
    const unsigned blocks = (unsigned)std::min<int64_t>((ntile + 3) / 4, kDiagWaves / 4);
    if (pk)
        hipLaunchKernelGGL(diag_r<true>, dim3(blocks), dim3(256), 0, s, r, cells, nslot, lay, ny, nz, nxp, part);
    else
        hipLaunchKernelGGL(diag_r<false>, dim3(blocks), dim3(256), 0, s, r, cells, nslot, lay, ny, nz, nxp, part);
    hipLaunchKernelGGL(diag_r_merge, dim3(1), dim3(256), 0, s, part, (int64_t)blocks * 4, out);
}

template <typename real>
void launch_slice(const real *F, const real *Ex, const real *Ey, const real *Ez, int emag, char axis, int pos,
                  int nx, int ny, int nz, int nxp, float *out, hipStream_t s) {
    int64_t n = axis == 'z' ? (int64_t)nx * ny : axis == 'x' ? (int64_t)ny * nz : (int64_t)nx * nz;
    hipLaunchKernelGGL(slice_k<real>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, F, Ex, Ey, Ez, emag,
                       axis, pos, nx, ny, nz, nxp, out);
}

template <typename real>
void launch_derive(const DeriveArgs &a, hipStream_t s) {
    const int64_t n = (int64_t)(a.i1 - a.i0) * (a.j1 - a.j0) * (a.k1 - a.k0);
    if (n <= 0) return;
    const int64_t nb = (n + 255) / 256;
    hipLaunchKernelGGL(derive_d<real>, dim3((unsigned)(nb < 8192 ? nb : 8192)), dim3(256), 0, s, a);
}

template <typename real>
void launch_dft(const DftArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(dft_accum<real>, dim3((unsigned)((a.nplane + 255) / 256)), dim3(256), 0, s, a);
}

template void launch_step_end<double>(const StepEndArgs &, hipStream_t);
template void launch_step_end<float>(const StepEndArgs &, hipStream_t);
template void launch_step_end_batch<double>(const StepEndArgs *, int, hipStream_t);
template void launch_step_end_batch<float>(const StepEndArgs *, int, hipStream_t);
template void launch_fill<double>(double *, double, int64_t, hipStream_t);
template void launch_fill<float>(float *, float, int64_t, hipStream_t);
template void launch_fill<unsigned long long>(unsigned long long *, unsigned long long, int64_t, hipStream_t);
template void launch_fill<int64_t>(int64_t *, int64_t, int64_t, hipStream_t);
template void launch_fill<uint8_t>(uint8_t *, uint8_t, int64_t, hipStream_t);
template void launch_reduce_e<double>(const double *, const double *, const double *, int, int, int, int,
                                      unsigned long long *, hipStream_t);
template void launch_reduce_e<float>(const float *, const float *, const float *, int, int, int, int,
                                     unsigned long long *, hipStream_t);
template void launch_derive<float>(const DeriveArgs &, hipStream_t);
template void launch_dft<double>(const DftArgs &, hipStream_t);
template void launch_dft<float>(const DftArgs &, hipStream_t);
template void launch_slice<double>(const double *, const double *, const double *, const double *, int, char, int,
                                   int, int, int, int, float *, hipStream_t);
template void launch_slice<float>(const float *, const float *, const float *, const float *, int, char, int, int,
                                  int, int, int, float *, hipStream_t);

}  // namespace phpfdtd
