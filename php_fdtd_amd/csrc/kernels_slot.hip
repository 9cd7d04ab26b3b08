// kernels_slot.hip - E update of ADE and lattice cells (fdtd.c:3404-3531).
//
// Cells of a material that carries ADE polarisation or a phonon lattice are
// listed once, in x-fastest order, as the slots of that material; the slot
// kernels run after the D/E kernel has produced this step's D there.
// Everything is done in the crystal frame with host-precomputed rotation R:
//   D_cr = R D,  D' = D_cr - P_ion,  P_ion = QeZV r^n           (3417-3427)
//   ADE:  P^{n+1} = Gp^-1 (2 P^n - Gm P^{n-1} + eps0 dt^2 wp2 R E^n)
//         E_cr = (1-lt) (D' - P^{n+1}) / (eps0 EPS_INF)         (3463-3479)
//   else: E_cr = (1-lt) lf Re(eps^-1) D' / eps0                   (3457-3461)
//   a = QeZM^T E_cr - DMM r^n;  r += v dt + a dt^2/2;  v += a dt  (3505-3528)
//   E = R^T E_cr                                                  (3489-3503)
//
// Lattice state layout (IonLayout): tiles of 64 slots; the plain form keeps
// r[tile][dof][64], the precision form the MFMA operand orders, so every
// wavefront access is 512 contiguous bytes streamed non-temporally.
//
// Plain form (FP64): one thread per slot, DMM r^n on the VALU in FP64 with
// DMM/QeZV/QeZM as wavefront-uniform (scalar) operands.
// Precision form (FP32): one wavefront per 64-slot tile; [DMM; QeZV] r^n is
// one v_mfma_f32_16x16x4f32 contraction per 16-slot group (A = the stacked
// 3N+3 x 3N matrix, pre-arranged in lane order on the host; B = r^n rounded
// to FP32), accumulated in FP32; P_ion falls out of the same product.  The
// r, v integration stays FP64.
#include "phpfdtd_internal.h"

namespace phpfdtd {

template <typename real>
__device__ __forceinline__ void rot(const real *R, const real x[3], real y[3]) {
    y[0] = R[0] * x[0] + R[1] * x[1] + R[2] * x[2];
    y[1] = R[3] * x[0] + R[4] * x[1] + R[5] * x[2];
    y[2] = R[6] * x[0] + R[7] * x[1] + R[8] * x[2];
}

template <typename real>
__device__ __forceinline__ void rot_t(const real *R, const real x[3], real y[3]) {
    y[0] = R[0] * x[0] + R[3] * x[1] + R[6] * x[2];
    y[1] = R[1] * x[0] + R[4] * x[1] + R[7] * x[2];
    y[2] = R[2] * x[0] + R[5] * x[1] + R[8] * x[2];
}

template <typename T>
__device__ __forceinline__ T ldnt(const T *p) {
    return __builtin_nontemporal_load(p);
}

template <typename T>
__device__ __forceinline__ void stnt(T v, T *p) {
    __builtin_nontemporal_store(v, p);
}

// ADE recurrence in the crystal frame for one slot; returns P^{n+1}.
template <typename real>
__device__ __forceinline__ void ade_step(const SlotArgs<real> &a, int64_t s, const real Eo[3], real Pn[3]) {
    const int64_t n = a.nslot;
    real P[3], Po[3], tv[3];
    for (int e = 0; e < 3; ++e) {
        P[e] = a.P[e * n + s];
        Po[e] = a.Po[e * n + s];
    }
    for (int e = 0; e < 3; ++e)
        tv[e] = 2 * P[e] - (a.Gm[e * 3] * Po[0] + a.Gm[e * 3 + 1] * Po[1] + a.Gm[e * 3 + 2] * Po[2]) +
                (a.W[e * 3] * Eo[0] + a.W[e * 3 + 1] * Eo[1] + a.W[e * 3 + 2] * Eo[2]);
    for (int e = 0; e < 3; ++e) Pn[e] = a.Gpi[e * 3] * tv[0] + a.Gpi[e * 3 + 1] * tv[1] + a.Gpi[e * 3 + 2] * tv[2];
    for (int e = 0; e < 3; ++e) {
        a.Po[e * n + s] = P[e];
        a.P[e * n + s] = Pn[e];
    }
}

// ADE-only materials (natom == 0): slot s.
template <typename real>
__device__ __forceinline__ void ade_slot(const SlotArgs<real> &a, const int64_t s) {
    if (s >= a.nslot) return;
    const size_t c = a.cells[s];
    const real Dl[3] = {a.Dx[c], a.Dy[c], a.Dz[c]};
    const real El[3] = {a.Ex[c], a.Ey[c], a.Ez[c]};
    real Dc[3], Ec[3], Pn[3], En[3], Eo[3];
    rot(a.R, Dl, Dc);
    rot(a.R, El, Ec);
    ade_step(a, s, Ec, Pn);
    for (int e = 0; e < 3; ++e) Dc[e] -= Pn[e];
    rot(a.A, Dc, En);
    rot_t(a.R, En, Eo);
    a.Ex[c] = Eo[0];
    a.Ey[c] = Eo[1];
    a.Ez[c] = Eo[2];
}

// Plain form: one thread per slot, everything FP64.  NA > 0 fixes natom.
template <int NA>
__device__ __forceinline__ void ion_f64_slot(const SlotArgs<double> &a, const double *__restrict__ dmm,
                                             const double *__restrict__ qezv, const double *__restrict__ qezm,
                                             double *__restrict__ rr, double *__restrict__ vv, const int64_t s) {
    constexpr int NMAX = NA > 0 ? 3 * NA : 3 * kMaxAtom;
    if (s >= a.nslot) return;
    const int N3 = NA > 0 ? 3 * NA : 3 * a.natom;
    const size_t c = a.cells[s];
    const int64_t tb = a.lay.r_index(s, 0);  // dof q at tb + 64 q (plain layout, same for v)
    const double Dl[3] = {a.Dx[c], a.Dy[c], a.Dz[c]};
    double r[NMAX];
#pragma unroll
    for (int q = 0; q < NMAX; ++q)
        if (q < N3) r[q] = ldnt(rr + tb + q * 64);
    double Dc[3];
    rot(a.R, Dl, Dc);
    double P[3] = {0, 0, 0};
#pragma unroll
    for (int q = 0; q < NMAX; ++q) {
        if (q < N3) {
            const int at = q / 3, d = q % 3;
            P[0] += sld(qezv, at * 9 + 0 + d) * r[q];
            P[1] += sld(qezv, at * 9 + 3 + d) * r[q];
            P[2] += sld(qezv, at * 9 + 6 + d) * r[q];
        }
    }
    double Dp[3] = {Dc[0] - P[0], Dc[1] - P[1], Dc[2] - P[2]};
    if (a.ade) {
        const double El[3] = {a.Ex[c], a.Ey[c], a.Ez[c]};
        double Eo[3], Pn[3];
        rot(a.R, El, Eo);
        ade_step(a, s, Eo, Pn);
        for (int e = 0; e < 3; ++e) Dp[e] -= Pn[e];
    }
    double Ec[3], El[3];
    rot(a.A, Dp, Ec);
    rot_t(a.R, Ec, El);
    a.Ex[c] = El[0];
    a.Ey[c] = El[1];
    a.Ez[c] = El[2];
    const double dt = a.dt, hdt2 = a.half_dt2;
#pragma unroll
    for (int q = 0; q < NMAX; ++q) {
        if (q < N3) {
            const int at = q / 3, d = q % 3;
            double acc = sld(qezm, at * 9 + 0 + d) * Ec[0] + sld(qezm, at * 9 + 3 + d) * Ec[1] +
                         sld(qezm, at * 9 + 6 + d) * Ec[2];
#pragma unroll
            for (int jq = 0; jq < NMAX; ++jq)
                if (jq < N3) acc -= sld(dmm, q * N3 + jq) * r[jq];
            const double vq = ldnt(vv + tb + q * 64);
            stnt(r[q] + (vq * dt + acc * hdt2), rr + tb + q * 64);
            stnt(vq + acc * dt, vv + tb + q * 64);
        }
    }
}

typedef float f32x4 __attribute__((ext_vector_type(4)));

// Precision form: one wavefront per 64-slot tile, on FP32 MFMA.  Operand
// layouts of v_mfma_f32_16x16x4f32 (lane l, gfx950):
//   A[m][k] in lane m + 16k,  B[k][n] in lane n + 16k,
//   D[m][n] in lane n + 16 (m / 4), element m % 4.
// r and v are stored in B-operand order (IonLayout, mfma = 1), so lane l of
// k-tile kt of group g owns dof 4 kt + l / 16 of slot 16 g + l % 16 and every
// global access of the kernel is one 512-byte row.  Per 16-slot group g:
//   1. acc = [-DMM; 0; QeZV] . r^n over the k-tiles (B = r^n rounded to
//      FP32); the QeZV rows at PR = roundup(3N, 4) give P_ion in elements
//      0..2 of lane group (PR % 16) / 4 of m-tile PR / 16.
//   2. lane 16g + n (which prefetched D of slot 16g + n) forms E_cr of its
//      slot; E_cr is spread as one more k-tile (E component k in lane
//      n + 16k) and acc += QeZM^T . E_cr, one MFMA per m-tile, so
//      acc = QeZM^T E_cr - DMM r^n = a (accumulator order).
//   3. a is transposed to B order through a 17-float-pitch LDS tile of the
//      wavefront, and each lane integrates its own dofs in FP64 from the r^n
//      it still holds and the v it prefetched.
// The groups are software-pipelined: r^n of group g+1 and v of group g are
// requested before group g's MFMA chain.
// aop[(mt*nk + kt)*64 + l] and qop[mt*64 + l] hold the A operands in lane order.
template <int NA>
__device__ __forceinline__ void ion_mfma_block(const SlotArgs<float> &a, const float *__restrict__ aop,
                                               const float *__restrict__ qop, double *__restrict__ rr,
                                               double *__restrict__ vv, const int64_t blk) {
    constexpr int NMX = ((3 * NA + 3) / 4 * 4 + 3 + 15) / 16;
    constexpr int NKX = (3 * NA + 3) / 4;
    constexpr int NAX = (3 * NA + 15) / 16;
    const int lane = threadIdx.x & 63;
    const int64_t tile = blk * 4 + (threadIdx.x >> 6);
    const int64_t ntile = (a.nslot + 63) >> 6;
    if (tile >= ntile) return;  // wave-uniform
    const IonLayout L = a.lay;
    // exact instantiations (natom == NA) use compile-time tile counts, so every
    // operand array is fully unrolled into registers; the generic one (NA =
    // kMaxAtom) runs to the material's own counts
    constexpr bool exact = NA != kMaxAtom;
    const int N3 = exact ? 3 * NA : L.n3;
    const int nk = exact ? NKX : L.nk, nma = exact ? NAX : L.nma;
    const int nm = exact ? NMX : ((N3 + 3) / 4 * 4 + 3 + 15) / 16;
    const int PR = (N3 + 3) & ~3;
    const int pm = PR >> 4, pg = (PR & 15) >> 2;
    const int grp = lane >> 4, col = lane & 15;
    const int64_t base = tile * L.tile_r();
    extern __shared__ float lds_a[];
    float *at = lds_a + (threadIdx.x >> 6) * (NKX * 4 * 17);  // this wavefront's a[dof][17]

    float A[NMX][NKX], AQ[NAX];
#pragma unroll
    for (int mt = 0; mt < NMX; ++mt)
#pragma unroll
        for (int kt = 0; kt < NKX; ++kt) A[mt][kt] = (mt < nm && kt < nk) ? aop[(mt * nk + kt) * 64 + lane] : 0.f;
#pragma unroll
    for (int mt = 0; mt < NAX; ++mt) AQ[mt] = mt < nma ? qop[mt * 64 + lane] : 0.f;

    // lane l prefetches D (and E^n for ADE) of slot 64 tile + l
    const int64_t my_slot = tile * 64 + lane;
    const bool my_live = my_slot < a.nslot;
    size_t my_cell = 0;
    float Dl[3] = {0.f, 0.f, 0.f}, Eold[3] = {0.f, 0.f, 0.f};
    if (my_live) {
        my_cell = a.cells[my_slot];
        Dl[0] = a.Dx[my_cell];
        Dl[1] = a.Dy[my_cell];
        Dl[2] = a.Dz[my_cell];
        if (a.ade) {
            Eold[0] = a.Ex[my_cell];
            Eold[1] = a.Ey[my_cell];
            Eold[2] = a.Ez[my_cell];
        }
    }
    const double dt = a.dt, hdt2 = a.half_dt2;
    double rnext[NKX];
#pragma unroll
    for (int kt = 0; kt < NKX; ++kt) rnext[kt] = kt < nk ? ldnt(rr + base + (int64_t)kt * 64 + lane) : 0.0;
#pragma unroll 1
    for (int g = 0; g < 4; ++g) {
        double rcur[NKX], vcur[NKX];
#pragma unroll
        for (int kt = 0; kt < NKX; ++kt) {
            rcur[kt] = rnext[kt];
            vcur[kt] = kt < nk ? ldnt(vv + base + ((int64_t)g * nk + kt) * 64 + lane) : 0.0;
        }
        if (g < 3) {
#pragma unroll
            for (int kt = 0; kt < NKX; ++kt)
                rnext[kt] = kt < nk ? ldnt(rr + base + ((int64_t)(g + 1) * nk + kt) * 64 + lane) : 0.0;
        }
        f32x4 acc[NMX];
#pragma unroll
        for (int mt = 0; mt < NMX; ++mt) acc[mt] = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kt = 0; kt < NKX; ++kt) {
            if (kt >= nk) break;
            const float b = (float)rcur[kt];
#pragma unroll
            for (int mt = 0; mt < NMX; ++mt)
                if (mt < nm) acc[mt] = __builtin_amdgcn_mfma_f32_16x16x4f32(A[mt][kt], b, acc[mt], 0, 0, 0);
        }
        // P_ion of slot 16g + col to lane 16g + col
        float p0 = 0.f, p1 = 0.f, p2 = 0.f;
#pragma unroll
        for (int mt = 0; mt < NMX; ++mt)
            if (mt == pm) { p0 = acc[mt][0]; p1 = acc[mt][1]; p2 = acc[mt][2]; }
        const int psrc = pg * 16 + col;
        p0 = __shfl(p0, psrc, 64);
        p1 = __shfl(p1, psrc, 64);
        p2 = __shfl(p2, psrc, 64);
        float Ec[3] = {0.f, 0.f, 0.f};
        if (grp == g && my_live) {
            float Dc[3];
            rot(a.R, Dl, Dc);
            float Dp[3] = {Dc[0] - p0, Dc[1] - p1, Dc[2] - p2};
            if (a.ade) {
                float Eo[3], Pn[3];
                rot(a.R, Eold, Eo);
                ade_step(a, my_slot, Eo, Pn);
                for (int e = 0; e < 3; ++e) Dp[e] -= Pn[e];
            }
            rot(a.A, Dp, Ec);
            float El[3];
            rot_t(a.R, Ec, El);
            a.Ex[my_cell] = El[0];
            a.Ey[my_cell] = El[1];
            a.Ez[my_cell] = El[2];
        }
        // E_cr of slot 16g + col as the B operand: component k in lane col + 16k
        const int esrc = 16 * g + col;
        const float e0 = __shfl(Ec[0], esrc, 64), e1 = __shfl(Ec[1], esrc, 64), e2 = __shfl(Ec[2], esrc, 64);
        const float eb = grp == 0 ? e0 : grp == 1 ? e1 : grp == 2 ? e2 : 0.f;
#pragma unroll
        for (int mt = 0; mt < NAX; ++mt)
            if (mt < nma) acc[mt] = __builtin_amdgcn_mfma_f32_16x16x4f32(AQ[mt], eb, acc[mt], 0, 0, 0);
        // a from accumulator order to B order through LDS (same wavefront only)
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int mt = 0; mt < NAX; ++mt) {
            if (mt >= nma) break;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int m = 16 * mt + 4 * grp + i;
                if (m < N3) at[m * 17 + col] = acc[mt][i];
            }
        }
        __builtin_amdgcn_wave_barrier();
        const bool live = tile * 64 + 16 * g + col < a.nslot;
#pragma unroll
        for (int kt = 0; kt < NKX; ++kt) {
            if (kt >= nk) break;
            const int dof = 4 * kt + grp;
            if (dof >= N3 || !live) continue;
            const double am = (double)at[dof * 17 + col];
            const int64_t idx = base + ((int64_t)g * nk + kt) * 64 + lane;
            stnt(rcur[kt] + (vcur[kt] * dt + am * hdt2), rr + idx);
            stnt(vcur[kt] + am * dt, vv + idx);
        }
    }
}

template <typename real>
__global__ __launch_bounds__(256) void ade_update(SlotArgs<real> a) {
    ade_slot(a, (int64_t)blockIdx.x * 256 + threadIdx.x);
}

template <int NA>
__global__ __launch_bounds__(256) void ion_update_f64(SlotArgs<double> a, const double *__restrict__ dmm,
                                                      const double *__restrict__ qezv,
                                                      const double *__restrict__ qezm, double *__restrict__ rr,
                                                      double *__restrict__ vv) {
    ion_f64_slot<NA>(a, dmm, qezv, qezm, rr, vv, (int64_t)blockIdx.x * 256 + threadIdx.x);
}

template <int NA>
__global__ __launch_bounds__(256) void ion_update_mfma(SlotArgs<float> a, const float *__restrict__ aop,
                                                       const float *__restrict__ qop, double *__restrict__ rr,
                                                       double *__restrict__ vv) {
    ion_mfma_block<NA>(a, aop, qop, rr, vv, blockIdx.x);
}

// Batched launches over the slot work of several domains: work-group g is
// block g - start[map[g]] of domain map[g].
template <typename real>
__global__ __launch_bounds__(256) void ade_update_batch(BatchLaunch<SlotArgs<real>> b) {
    const int rid = (int)sld(b.map, blockIdx.x);
    const SlotArgs<real> a = batch_args(b, rid);
    ade_slot(a, (int64_t)(blockIdx.x - sld(b.start, rid)) * 256 + threadIdx.x);
}

template <int NA>
__global__ __launch_bounds__(256) void ion_update_f64_batch(BatchLaunch<SlotArgs<double>> b) {
    const int rid = (int)sld(b.map, blockIdx.x);
    const SlotArgs<double> a = batch_args(b, rid);
    ion_f64_slot<NA>(a, a.dmm, a.qezv, a.qezm, a.r, a.v, (int64_t)(blockIdx.x - sld(b.start, rid)) * 256 + threadIdx.x);
}

template <int NA>
__global__ __launch_bounds__(256) void ion_update_mfma_batch(BatchLaunch<SlotArgs<float>> b) {
    const int rid = (int)sld(b.map, blockIdx.x);
    const SlotArgs<float> a = batch_args(b, rid);
    ion_mfma_block<NA>(a, a.aop, a.qop, a.r, a.v, blockIdx.x - sld(b.start, rid));
}

template <typename real>
void launch_ade(const SlotArgs<real> &a, hipStream_t st) {
    const unsigned nb = (unsigned)((a.nslot + 255) / 256);
    if (nb) hipLaunchKernelGGL(ade_update<real>, dim3(nb), dim3(256), 0, st, a);
}

template <typename real>
void launch_ion(const SlotArgs<real> &a, hipStream_t st);

template <>
void launch_ion<double>(const SlotArgs<double> &a, hipStream_t st) {
    const unsigned nb = (unsigned)((a.nslot + 255) / 256);
    if (!nb) return;
#define PHPFDTD_ION_CASE(NN)                                                                              \
    case NN:                                                                                              \
        hipLaunchKernelGGL(ion_update_f64<NN>, dim3(nb), dim3(256), 0, st, a, a.dmm, a.qezv, a.qezm, a.r, \
                           a.v);                                                                          \
        return;
    switch (a.natom) {
        PHPFDTD_ION_CASE(1)
        PHPFDTD_ION_CASE(2)
        PHPFDTD_ION_CASE(3)
        PHPFDTD_ION_CASE(4)
        PHPFDTD_ION_CASE(6)
        PHPFDTD_ION_CASE(8)
        PHPFDTD_ION_CASE(12)
        PHPFDTD_ION_CASE(16)
        default:
            hipLaunchKernelGGL(ion_update_f64<0>, dim3(nb), dim3(256), 0, st, a, a.dmm, a.qezv, a.qezm, a.r, a.v);
    }
#undef PHPFDTD_ION_CASE
}

template <>
void launch_ion<float>(const SlotArgs<float> &a, hipStream_t st) {
    const int64_t ntile = (a.nslot + 63) / 64;
    const unsigned nb = (unsigned)((ntile + 3) / 4);
    if (!nb) return;
    // the host lays out aop/qop for the instantiation mfma_inst(natom)
    switch (mfma_inst(a.natom)) {
#define PHPFDTD_MFMA_CASE(NN)                                                                                \
    case NN:                                                                                                 \
        hipLaunchKernelGGL(ion_update_mfma<NN>, dim3(nb), dim3(256), (size_t)4 * mfma_nk(NN) * 4 * 17 * 4, st, a, \
                           a.aop, a.qop, a.r, a.v);                                                          \
        return;
        PHPFDTD_MFMA_CASE(1)
        PHPFDTD_MFMA_CASE(2)
        PHPFDTD_MFMA_CASE(3)
        PHPFDTD_MFMA_CASE(4)
        PHPFDTD_MFMA_CASE(6)
        PHPFDTD_MFMA_CASE(8)
        PHPFDTD_MFMA_CASE(12)
        PHPFDTD_MFMA_CASE(16)
        PHPFDTD_MFMA_CASE(kMaxAtom)
#undef PHPFDTD_MFMA_CASE
    }
}

// Work-groups of one domain's slot kernel (the unit of a batched launch).
template <typename real>
int64_t slot_blocks(const SlotArgs<real> &a) {
    if (a.natom > 0 && sizeof(real) == 4) return ((a.nslot + 63) / 64 + 3) / 4;
    return (a.nslot + 255) / 256;
}

// Instantiation key of a domain's slot kernel: 0 for ADE-only materials,
// else the natom template the kernel is compiled for.
template <typename real>
int slot_kind(const SlotArgs<real> &a) {
    if (a.natom == 0) return 0;
    if (sizeof(real) == 4) return mfma_inst(a.natom);
    switch (a.natom) {
        case 1: case 2: case 3: case 4: case 6: case 8: case 12: case 16: return a.natom;
        default: return kMaxAtom;
    }
}

template <typename real>
void launch_slot_batch(const BatchLaunch<SlotArgs<real>> &b, int kind, hipStream_t st);

template <>
void launch_slot_batch<double>(const BatchLaunch<SlotArgs<double>> &b, int kind, hipStream_t st) {
    if (b.total <= 0) return;
    const dim3 g(b.total), blk(256);
    switch (kind) {
        case 0: hipLaunchKernelGGL(ade_update_batch<double>, g, blk, 0, st, b); return;
#define PHPFDTD_ION_CASE(NN) \
    case NN: hipLaunchKernelGGL(ion_update_f64_batch<NN>, g, blk, 0, st, b); return;
        PHPFDTD_ION_CASE(1)
        PHPFDTD_ION_CASE(2)
        PHPFDTD_ION_CASE(3)
        PHPFDTD_ION_CASE(4)
        PHPFDTD_ION_CASE(6)
        PHPFDTD_ION_CASE(8)
        PHPFDTD_ION_CASE(12)
        PHPFDTD_ION_CASE(16)
#undef PHPFDTD_ION_CASE
        default: hipLaunchKernelGGL(ion_update_f64_batch<0>, g, blk, 0, st, b); return;
    }
}

template <>
void launch_slot_batch<float>(const BatchLaunch<SlotArgs<float>> &b, int kind, hipStream_t st) {
    if (b.total <= 0) return;
    const dim3 g(b.total), blk(256);
    switch (kind) {
        case 0: hipLaunchKernelGGL(ade_update_batch<float>, g, blk, 0, st, b); return;
#define PHPFDTD_MFMA_CASE(NN) \
    case NN: hipLaunchKernelGGL(ion_update_mfma_batch<NN>, g, blk, (size_t)4 * mfma_nk(NN) * 4 * 17 * 4, st, b); return;
        PHPFDTD_MFMA_CASE(1)
        PHPFDTD_MFMA_CASE(2)
        PHPFDTD_MFMA_CASE(3)
        PHPFDTD_MFMA_CASE(4)
        PHPFDTD_MFMA_CASE(6)
        PHPFDTD_MFMA_CASE(8)
        PHPFDTD_MFMA_CASE(12)
        PHPFDTD_MFMA_CASE(16)
        PHPFDTD_MFMA_CASE(kMaxAtom)
#undef PHPFDTD_MFMA_CASE
    }
}

template void launch_ade<double>(const SlotArgs<double> &, hipStream_t);
template void launch_ade<float>(const SlotArgs<float> &, hipStream_t);
template int64_t slot_blocks<double>(const SlotArgs<double> &);
template int64_t slot_blocks<float>(const SlotArgs<float> &);
template int slot_kind<double>(const SlotArgs<double> &);
template int slot_kind<float>(const SlotArgs<float> &);

}  // namespace phpfdtd
