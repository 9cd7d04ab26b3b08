// kernels_slot.hip - E update of ADE and lattice cells (fdtd.c:3404-3531).
//
// Cells of a material that carries ADE polarisation or a phonon lattice are
// listed once, in x-fastest order, as the slots of that material; one thread
// updates one slot after the D/E kernel has produced this step's D there.
// Everything is done in the crystal frame with host-precomputed rotation R:
//   D_cr = R D,  D' = D_cr - P_ion,  P_ion = QeZV r^n           (3417-3427)
//   ADE:  P^{n+1} = Gp^-1 (2 P^n - Gm P^{n-1} + eps0 dt^2 wp2 R E^n)
//         E_cr = (1-lt) (D' - P^{n+1}) / (eps0 EPS_INF)         (3463-3479)
//   else: E_cr = (1-lt) lf Re(eps^-1) D' / eps0                   (3457-3461)
//   a = QeZM^T E_cr - DMM r^n;  r += v dt + a dt^2/2;  v += a dt  (3505-3528)
//   E = R^T E_cr                                                  (3489-3503)
// The lattice contraction DMM r^n (3N x 3N against the slot's 3N-vector) is
// evaluated with wavefront-uniform (scalar-cache) operands: DMM, QeZV and
// QeZM are the same for every slot of the material.
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

// ADE-only materials (natom == 0).
template <typename real>
__global__ __launch_bounds__(256) void ade_update(SlotArgs<real> a) {
    const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (s >= a.nslot) return;
    const size_t c = a.cells[s];
    const int64_t n = a.nslot;
    const real Dl[3] = {a.Dx[c], a.Dy[c], a.Dz[c]};
    const real El[3] = {a.Ex[c], a.Ey[c], a.Ez[c]};
    real Dc[3], Ec[3];
    rot(a.R, Dl, Dc);
    rot(a.R, El, Ec);
    real P[3], Po[3], tv[3], Pn[3], En[3], Eo[3];
    for (int e = 0; e < 3; ++e) {
        P[e] = a.P[e * n + s];
        Po[e] = a.Po[e * n + s];
    }
    for (int e = 0; e < 3; ++e)
        tv[e] = 2 * P[e] - (a.Gm[e * 3] * Po[0] + a.Gm[e * 3 + 1] * Po[1] + a.Gm[e * 3 + 2] * Po[2]) +
                (a.W[e * 3] * Ec[0] + a.W[e * 3 + 1] * Ec[1] + a.W[e * 3 + 2] * Ec[2]);
    for (int e = 0; e < 3; ++e) Pn[e] = a.Gpi[e * 3] * tv[0] + a.Gpi[e * 3 + 1] * tv[1] + a.Gpi[e * 3 + 2] * tv[2];
    for (int e = 0; e < 3; ++e) Dc[e] -= Pn[e];
    rot(a.A, Dc, En);
    rot_t(a.R, En, Eo);
    a.Ex[c] = Eo[0];
    a.Ey[c] = Eo[1];
    a.Ez[c] = Eo[2];
    for (int e = 0; e < 3; ++e) {
        a.Po[e * n + s] = P[e];
        a.P[e * n + s] = Pn[e];
    }
}

// Lattice materials.  acc_t is the type of the DMM/QeZV/QeZM contraction
// (double in the plain form, float in the precision form); r, v are FP64.
// NA > 0 fixes natom at compile time (fully unrolled, r held in registers).
template <typename real, typename acc_t, int NA>
__global__ __launch_bounds__(256) void ion_update(SlotArgs<real> a, const acc_t *__restrict__ dmm,
                                                  const acc_t *__restrict__ qezv,
                                                  const acc_t *__restrict__ qezm, double *__restrict__ rr,
                                                  double *__restrict__ vv) {
    constexpr int NMAX = NA > 0 ? 3 * NA : 3 * kMaxAtom;
    const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (s >= a.nslot) return;
    const int N3 = NA > 0 ? 3 * NA : 3 * a.natom;
    const size_t c = a.cells[s];
    const int64_t n = a.nslot;
    const real Dl[3] = {a.Dx[c], a.Dy[c], a.Dz[c]};
    real Dc[3];
    rot(a.R, Dl, Dc);

    double r[NMAX];
    acc_t ra[NMAX];
#pragma unroll
    for (int q = 0; q < NMAX; ++q) {
        if (q < N3) {
            r[q] = rr[q * n + s];
            ra[q] = (acc_t)r[q];
        }
    }
    // P_ion = QeZV . r  (QeZV[atom][E comp][dir])
    acc_t P[3] = {0, 0, 0};
#pragma unroll
    for (int q = 0; q < NMAX; ++q) {
        if (q < N3) {
            const int at = q / 3, d = q % 3;
            P[0] += qezv[at * 9 + 0 + d] * ra[q];
            P[1] += qezv[at * 9 + 3 + d] * ra[q];
            P[2] += qezv[at * 9 + 6 + d] * ra[q];
        }
    }
    real Dp[3] = {Dc[0] - (real)P[0], Dc[1] - (real)P[1], Dc[2] - (real)P[2]};
    real Ec[3];
    if (a.ade) {
        const real El[3] = {a.Ex[c], a.Ey[c], a.Ez[c]};
        real Eo[3], Pp[3], Po[3], tv[3], Pn[3];
        rot(a.R, El, Eo);
        for (int e = 0; e < 3; ++e) {
            Pp[e] = a.P[e * n + s];
            Po[e] = a.Po[e * n + s];
        }
        for (int e = 0; e < 3; ++e)
            tv[e] = 2 * Pp[e] - (a.Gm[e * 3] * Po[0] + a.Gm[e * 3 + 1] * Po[1] + a.Gm[e * 3 + 2] * Po[2]) +
                    (a.W[e * 3] * Eo[0] + a.W[e * 3 + 1] * Eo[1] + a.W[e * 3 + 2] * Eo[2]);
        for (int e = 0; e < 3; ++e)
            Pn[e] = a.Gpi[e * 3] * tv[0] + a.Gpi[e * 3 + 1] * tv[1] + a.Gpi[e * 3 + 2] * tv[2];
        for (int e = 0; e < 3; ++e) {
            a.Po[e * n + s] = Pp[e];
            a.P[e * n + s] = Pn[e];
            Dp[e] -= Pn[e];
        }
    }
    rot(a.A, Dp, Ec);
    {
        real El[3];
        rot_t(a.R, Ec, El);
        a.Ex[c] = El[0];
        a.Ey[c] = El[1];
        a.Ez[c] = El[2];
    }
    const acc_t E0 = (acc_t)Ec[0], E1 = (acc_t)Ec[1], E2 = (acc_t)Ec[2];
    const double dt = a.dt, hdt2 = a.half_dt2;
#pragma unroll
    for (int q = 0; q < NMAX; ++q) {
        if (q < N3) {
            const int at = q / 3, d = q % 3;
            acc_t acc = qezm[at * 9 + 0 + d] * E0 + qezm[at * 9 + 3 + d] * E1 + qezm[at * 9 + 6 + d] * E2;
#pragma unroll
            for (int jq = 0; jq < NMAX; ++jq)
                if (jq < N3) acc -= dmm[q * N3 + jq] * ra[jq];
            const double aq = (double)acc;
            const double vq = vv[q * n + s];
            rr[q * n + s] = r[q] + (vq * dt + aq * hdt2);
            vv[q * n + s] = vq + aq * dt;
        }
    }
}

template <typename real>
void launch_ade(const SlotArgs<real> &a, hipStream_t st) {
    const unsigned nb = (unsigned)((a.nslot + 255) / 256);
    if (nb) hipLaunchKernelGGL(ade_update<real>, dim3(nb), dim3(256), 0, st, a);
}

template <typename real, typename acc_t>
static void launch_ion_t(const SlotArgs<real> &a, const acc_t *dmm, const acc_t *qezv, const acc_t *qezm,
                         hipStream_t st) {
    const unsigned nb = (unsigned)((a.nslot + 255) / 256);
    if (!nb) return;
#define PHPFDTD_ION_CASE(NN)                                                                      \
    case NN:                                                                                      \
        hipLaunchKernelGGL((ion_update<real, acc_t, NN>), dim3(nb), dim3(256), 0, st, a, dmm, qezv, \
                           qezm, a.r, a.v);                                                       \
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
            hipLaunchKernelGGL((ion_update<real, acc_t, 0>), dim3(nb), dim3(256), 0, st, a, dmm, qezv, qezm,
                               a.r, a.v);
    }
#undef PHPFDTD_ION_CASE
}

template <typename real>
void launch_ion(const SlotArgs<real> &a, hipStream_t st);

template <>
void launch_ion<double>(const SlotArgs<double> &a, hipStream_t st) {
    launch_ion_t<double, double>(a, a.dmm, a.qezv, a.qezm, st);
}

template <>
void launch_ion<float>(const SlotArgs<float> &a, hipStream_t st) {
    launch_ion_t<float, float>(a, a.dmm32, a.qezv32, a.qezm32, st);
}

template void launch_ade<double>(const SlotArgs<double> &, hipStream_t);
template void launch_ade<float>(const SlotArgs<float> &, hipStream_t);

}  // namespace phpfdtd
