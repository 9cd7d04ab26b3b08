// kernels_aux.hip - per-step epilogue, diagnostics and sampling kernels.
#include "phpfdtd_internal.h"

namespace phpfdtd {

// Per-step epilogue (one thread): the monitor row of fdtd.c:3806-3851 for the
// step that just finished, then the device step counter advances.  The row is
// written at index `step`, so graph replays need no host bookkeeping.  part 0
// writes the whole row; part 1 only D and H (taken before the next step's H
// update starts), part 2 t, E, R, V and advances the step.
__device__ __forceinline__ void step_row(const StepEndArgs &a, bool dh, bool rest, const void *const *F, bool f64) {
    const int64_t st = *a.step;
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
                    rv[0] = pk_unpack(static_cast<const float *>(a.mon_r)[a.mon_ridx],
                                      static_cast<const int8_t *>(a.mon_r)[a.mon_rc]);
                    rv[1] = pk_unpack(static_cast<const float *>(a.mon_v)[a.mon_vidx],
                                      static_cast<const int8_t *>(a.mon_v)[a.mon_vc]);
                } else {
                    rv[0] = static_cast<const double *>(a.mon_r)[a.mon_ridx];
                    rv[1] = static_cast<const double *>(a.mon_v)[a.mon_vidx];
                }
            }
            row[10] = rv[0];
            row[11] = rv[1];
        }
        if (dh)
            for (int q = 3; q < 9; ++q) row[1 + q] = val(q);
    }
    if (rest) *a.step = st + 1;
}

template <typename real>
__global__ void step_end(StepEndArgs a) {
    if (threadIdx.x != 0) return;
    step_row(a, a.part != 2, a.part != 1, a.F, sizeof(real) == 8);
}

// step_end of every domain of a batch: work-group b handles domain b.
template <typename real>
__global__ void step_end_batch(const StepEndArgs *args, int n) {
    if (threadIdx.x != 0 || (int)blockIdx.x >= n) return;
    const StepEndArgs a = args[blockIdx.x];
    step_row(a, true, true, a.F, sizeof(real) == 8);
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
// non-negative, so their IEEE bit patterns order like the values.
template <typename real>
__global__ __launch_bounds__(256) void reduce_e(const real *Ex, const real *Ey, const real *Ez, int nx, int ny,
                                                int nz, int nxp, unsigned long long *out) {
    const int64_t total = (int64_t)nxp * ny * nz;
    unsigned long long me = 0, mx = 0;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < total;
         c += (int64_t)gridDim.x * blockDim.x) {
        if ((int)(c % nxp) >= nx) continue;
        const double ex = Ex[c], ey = Ey[c], ez = Ez[c];
        const double e = sqrt(ex * ex + ey * ey + ez * ez);
        const unsigned long long be = __double_as_longlong(e), bx = __double_as_longlong(fabs(ex));
        me = me > be ? me : be;
        mx = mx > bx ? mx : bx;
    }
    me = wave_max_u64(me);
    mx = wave_max_u64(mx);
    if ((threadIdx.x & 63) == 0) {
        atomicMax(out + 0, me);
        atomicMax(out + 1, mx);
    }
}

// max |r| over one lattice material's state (FP64 values, or the FP32 words
// of compensated state: max |r| rounded to FP32; the compensation rows, the
// elements of a tile from `words` on, are skipped).
template <typename T>
__global__ __launch_bounds__(256) void reduce_r(const T *r, int64_t n, int64_t tile, int64_t words,
                                                unsigned long long *out) {
    unsigned long long m = 0;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        if (q % tile >= words) continue;
        const unsigned long long b = __double_as_longlong(fabs((double)r[q]));
        m = m > b ? m : b;
    }
    m = wave_max_u64(m);
    if ((threadIdx.x & 63) == 0) atomicMax(out + 2, m);
}

// First argmax in the reference's loop order (i, j, k, atom, dir): among the
// dofs equal to the maximum, the smallest key ((i*ny + j)*nz + k)*96 + q.
// r is in the material's lattice layout (IonLayout), values as reduce_r reads them.
template <typename T>
__global__ __launch_bounds__(256) void argmax_r(const T *r, const uint32_t *cells, int64_t nslot, IonLayout lay,
                                                int ny, int nz, int nxp, unsigned long long *out) {
    const unsigned long long target = out[2];
    const int64_t total = (nslot + 63) / 64 * lay.tile_r();
    unsigned long long best = ~0ull;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total;
         q += (int64_t)gridDim.x * blockDim.x) {
        int64_t s;
        int dof;
        if (!lay.r_decode(q, nslot, &s, &dof)) continue;  // padding, compensation words
        if ((unsigned long long)__double_as_longlong(fabs((double)r[q])) != target) continue;
        const int64_t c = cells[s];
        const int64_t i = c % nxp, j = (c / nxp) % ny, k = c / ((int64_t)nxp * ny);
        const unsigned long long key = (unsigned long long)(((i * ny + j) * nz + k) * 96 + dof);
        best = best < key ? best : key;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(best, off, 64);
        best = best < o ? best : o;
    }
    if ((threadIdx.x & 63) == 0) atomicMin(out + 3, best);
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
// before the step counter advances, so it sees the step just updated.
template <typename real>
__global__ __launch_bounds__(256) void dft_accum(DftArgs a) {
    const int64_t st = *a.step - a.t0;
    if (st < 0 || st % a.interval) return;
    const int64_t n = st / a.interval;
    if (n >= a.nsamp) return;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t c = plane_cell(q, a.axis, a.pos, a.nx, a.ny, a.nz, a.nxp);
    if (c < 0) return;
    const real *F = static_cast<const real *>(a.F);
    const real *Ex = static_cast<const real *>(a.Ex), *Ey = static_cast<const real *>(a.Ey),
               *Ez = static_cast<const real *>(a.Ez);
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
    hipLaunchKernelGGL(reduce_e<real>, dim3(2048), dim3(256), 0, s, Ex, Ey, Ez, nx, ny, nz, nxp, out);
}

// r: ntile tiles of layout lay; double values, or float words when f32 (compensated state).
void launch_reduce_r(const void *r, bool f32, int64_t ntile, IonLayout lay, unsigned long long *out, hipStream_t s) {
    const int64_t tile = lay.tile_r(), n = ntile * tile;
    if (n <= 0) return;
    if (f32)
        hipLaunchKernelGGL(reduce_r<float>, dim3(1024), dim3(256), 0, s, static_cast<const float *>(r), n, tile,
                           (int64_t)kPkWords * lay.ng(), out);
    else
        hipLaunchKernelGGL(reduce_r<double>, dim3(1024), dim3(256), 0, s, static_cast<const double *>(r), n, tile,
                           tile, out);
}

void launch_argmax_r(const void *r, bool f32, const uint32_t *cells, int64_t nslot, IonLayout lay, int ny, int nz,
                     int nxp, unsigned long long *out, hipStream_t s) {
    if (nslot <= 0) return;
    if (f32)
        hipLaunchKernelGGL(argmax_r<float>, dim3(1024), dim3(256), 0, s, static_cast<const float *>(r), cells, nslot,
                           lay, ny, nz, nxp, out);
    else
        hipLaunchKernelGGL(argmax_r<double>, dim3(1024), dim3(256), 0, s, static_cast<const double *>(r), cells,
                           nslot, lay, ny, nz, nxp, out);
}

template <typename real>
void launch_slice(const real *F, const real *Ex, const real *Ey, const real *Ez, int emag, char axis, int pos,
                  int nx, int ny, int nz, int nxp, float *out, hipStream_t s) {
    int64_t n = axis == 'z' ? (int64_t)nx * ny : axis == 'x' ? (int64_t)ny * nz : (int64_t)nx * nz;
    hipLaunchKernelGGL(slice_k<real>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, F, Ex, Ey, Ez, emag,
                       axis, pos, nx, ny, nz, nxp, out);
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
template void launch_dft<double>(const DftArgs &, hipStream_t);
template void launch_dft<float>(const DftArgs &, hipStream_t);
template void launch_slice<double>(const double *, const double *, const double *, const double *, int, char, int,
                                   int, int, int, int, float *, hipStream_t);
template void launch_slice<float>(const float *, const float *, const float *, const float *, int, char, int, int,
                                  int, int, int, float *, hipStream_t);

}  // namespace phpfdtd
