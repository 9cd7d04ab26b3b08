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
// Lattice state layout (IonLayout): tiles of 64 slots, a lane owns a slot and
// moves a group of its dofs as one 16-byte vector, so every wavefront access
// is 1 KiB contiguous, streamed non-temporally: FP64 values in dof pairs, or
// compensated FP32 (an FP32 word and a 4-bit correction per value) in dof
// quads, the tile's correction codes in 1 KiB rows after its word rows (eight
// groups' codes per 16-byte lane vector).
//
// Plain form (FP64): one thread per slot, DMM r^n on the VALU in FP64 with
// DMM/QeZV/QeZM as wavefront-uniform (scalar) operands.
// Precision form (FP32): the same slot-per-lane streams; -DMM r^n runs on FP32
// MFMA (v_mfma_f32_16x16x4f32) after an LDS transpose of r^n rounded to FP32,
// accumulated in FP32; P_ion and QeZM^T E_cr on the VALU in FP32.  The r, v
// integration stays FP64; r, v are stored compensated (PK) or in FP64.
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

typedef double f64x2 __attribute__((ext_vector_type(2)));

// Dof pair p of the slot at lane offset `base` (IonLayout: pairs 128 doubles apart).
__device__ __forceinline__ f64x2 ldpair(const double *base, int p) {
    return __builtin_nontemporal_load(reinterpret_cast<const f64x2 *>(base + (int64_t)p * 128));
}
__device__ __forceinline__ void stpair(f64x2 v, double *base, int p) {
    __builtin_nontemporal_store(v, reinterpret_cast<f64x2 *>(base + (int64_t)p * 128));
}

// E sets of this step (double-buffered, YeeArgs): E^n is read from set
// step & 1, the new E written to the other.
template <typename real>
struct SlotE {
    const real *o[3];
    real *n[3];
};
template <typename real>
__device__ __forceinline__ SlotE<real> slot_e(const SlotArgs<real> &a) {
    const bool odd = sld(a.step, 0) & 1;
    SlotE<real> e;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        e.o[q] = odd ? a.E[1][q] : a.E[0][q];
        e.n[q] = odd ? a.E[0][q] : a.E[1][q];
    }
    return e;
}

// ADE recurrence in the crystal frame for one slot; returns P^{n+1}.
// FP64: P^{n+1} = Gpi (2 P - Gm P^{n-1} + W E) on (P, P^{n-1}), the
// reference's order (fdtd.c:3463-3479).  FP32: the same recurrence in
// increment form on (P, J = P - P^{n-1}) - the P^n terms cancel exactly -
// u = J - Ga J + W E, J' = u - Gb u, P' = P + J': every correction is a small
// term added to its O(1) value, so the marginal mode of P (the Drude
// integrator, eigenvalue 1) stays exact in FP32, where the two-level form
// rounds Gm = I - Ga and Gpi to 24 bits and drifts.
template <typename real>
__device__ __forceinline__ void ade_step(const SlotArgs<real> &a, int64_t s, const real Eo[3], real Pn[3]) {
    const int64_t n = a.nslot;
    real P[3], Po[3], tv[3];
    for (int e = 0; e < 3; ++e) {
        P[e] = a.P[e * n + s];
        Po[e] = a.Po[e * n + s];  // P^{n-1} (FP64) or J (FP32)
    }
    if constexpr (sizeof(real) == 4) {
        real u[3], J[3];
        for (int e = 0; e < 3; ++e)
            u[e] = Po[e] - (a.Ga[e * 3] * Po[0] + a.Ga[e * 3 + 1] * Po[1] + a.Ga[e * 3 + 2] * Po[2]) +
                   (a.W[e * 3] * Eo[0] + a.W[e * 3 + 1] * Eo[1] + a.W[e * 3 + 2] * Eo[2]);
        for (int e = 0; e < 3; ++e) J[e] = u[e] - (a.Gb[e * 3] * u[0] + a.Gb[e * 3 + 1] * u[1] + a.Gb[e * 3 + 2] * u[2]);
        for (int e = 0; e < 3; ++e) {
            Pn[e] = P[e] + J[e];
            a.Po[e * n + s] = J[e];
            a.P[e * n + s] = Pn[e];
        }
    } else {
        for (int e = 0; e < 3; ++e)
            tv[e] = 2 * P[e] - (a.Gm[e * 3] * Po[0] + a.Gm[e * 3 + 1] * Po[1] + a.Gm[e * 3 + 2] * Po[2]) +
                    (a.W[e * 3] * Eo[0] + a.W[e * 3 + 1] * Eo[1] + a.W[e * 3 + 2] * Eo[2]);
        for (int e = 0; e < 3; ++e) Pn[e] = a.Gpi[e * 3] * tv[0] + a.Gpi[e * 3 + 1] * tv[1] + a.Gpi[e * 3 + 2] * tv[2];
        for (int e = 0; e < 3; ++e) {
            a.Po[e * n + s] = P[e];
            a.P[e * n + s] = Pn[e];
        }
    }
}

// ADE-only materials (natom == 0): slot s.
template <typename real>
__device__ __forceinline__ void ade_slot(const SlotArgs<real> &a, const int64_t s) {
    if (s >= a.nslot) return;
    const size_t c = a.cells[s];
    const SlotE<real> E = slot_e(a);
    const real Dl[3] = {a.Dx[c], a.Dy[c], a.Dz[c]};
    const real El[3] = {E.o[0][c], E.o[1][c], E.o[2][c]};
    real Dc[3], Ec[3], Pn[3], En[3], Eo[3];
    rot(a.R, Dl, Dc);
    rot(a.R, El, Ec);
    ade_step(a, s, Ec, Pn);
    for (int e = 0; e < 3; ++e) Dc[e] -= Pn[e];
    rot(a.A, Dc, En);
    rot_t(a.R, En, Eo);
    E.n[0][c] = Eo[0];
    E.n[1][c] = Eo[1];
    E.n[2][c] = Eo[2];
}

// Plain form: one thread per slot, everything FP64.  NA > 0 fixes natom.
template <int NA>
__device__ __forceinline__ void ion_f64_slot(const SlotArgs<double> &a, const double *__restrict__ dmm,
                                             const double *__restrict__ qezv, const double *__restrict__ qezm,
                                             double *__restrict__ rr, double *__restrict__ vv, const int64_t s) {
    constexpr int NMAX = NA > 0 ? (3 * NA + 1) & ~1 : 3 * kMaxAtom;  // even: whole dof pairs
    if (s >= a.nslot) return;
    const int N3 = NA > 0 ? 3 * NA : 3 * a.natom;
    const int NP = (N3 + 1) / 2;
    const size_t c = a.cells[s];
    const int64_t tb = a.lay.r_index(s, 0);  // dof pair p at tb + 128 p (same for v)
    const double Dl[3] = {a.Dx[c], a.Dy[c], a.Dz[c]};
    double r[NMAX];
#pragma unroll
    for (int p = 0; p < NMAX / 2; ++p)
        if (p < NP) {
            const f64x2 v2 = ldpair(rr + tb, p);
            r[2 * p] = v2.x;
            r[2 * p + 1] = v2.y;
        }
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
    const SlotE<double> E = slot_e(a);
    if (a.ade) {
        const double El[3] = {E.o[0][c], E.o[1][c], E.o[2][c]};
        double Eo[3], Pn[3];
        rot(a.R, El, Eo);
        ade_step(a, s, Eo, Pn);
        for (int e = 0; e < 3; ++e) Dp[e] -= Pn[e];
    }
    double Ec[3], El[3];
    rot(a.A, Dp, Ec);
    rot_t(a.R, Ec, El);
    E.n[0][c] = El[0];
    E.n[1][c] = El[1];
    E.n[2][c] = El[2];
    const double dt = a.dt, hdt2 = a.half_dt2;
#pragma unroll
    for (int p = 0; p < NMAX / 2; ++p) {
        if (p >= NP) break;
        f64x2 v2 = ldpair(vv + tb, p);
        f64x2 r2 = {r[2 * p], r[2 * p + 1]};
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int q = 2 * p + h;
            if (q >= N3) break;  // padding dof of an odd 3N stays 0
            const int at = q / 3, d = q % 3;
            double acc = sld(qezm, at * 9 + 0 + d) * Ec[0] + sld(qezm, at * 9 + 3 + d) * Ec[1] +
                         sld(qezm, at * 9 + 6 + d) * Ec[2];
#pragma unroll
            for (int jq = 0; jq < NMAX; ++jq)
                if (jq < N3) acc -= sld(dmm, q * N3 + jq) * r[jq];
            const double vq = v2[h];
            r2[h] = r[q] + (vq * dt + acc * hdt2);
            v2[h] = vq + acc * dt;
        }
        stpair(r2, rr + tb, p);
        stpair(v2, vv + tb, p);
    }
}

typedef float f32x4 __attribute__((ext_vector_type(4)));

// Precision form: lane l of a wavefront owns slot 64 tile + l and all 3N dofs
// of its r^n and v^n - the memory pattern of the plain form (dof groups as
// 16-byte vectors, every access one 1 KiB row, all of r^n requested up front).  Only the
// contraction -DMM . r^n runs on FP32 MFMA (v_mfma_f32_16x16x4f32; lane l on
// gfx950: A[m][k] in lane m + 16k, B[k][n] in lane n + 16k, D[m][n] in lane
// n + 16 (m / 4), element m % 4):
//   1. r^n is rounded to FP32 into LDS rows (dof q, 64 slots) of the
//      wavefront; P_ion = QeZV r^n on the VALU in FP32;
//   2. E_cr and the lab-frame E of the lane's slot, stored by the whole
//      wavefront (fdtd.c:3404-3503);
//   3. per 16-slot group g: acc_g = -DMM . r^n over the k-tiles, B read from
//      the LDS rows, A (-DMM in lane order) from an LDS table staged once per
//      work-group;
//   4. acc_g goes back to the LDS rows (dof m, slot 16g + n) and each lane
//      reads its own dofs: a = acc + QeZM^T E_cr (VALU, FP32), then the FP64
//      integration r += v dt + a dt^2/2, v += a dt (fdtd.c:3505-3528).
// QeZV / QeZM are read through the scalar cache from qop = [QeZV | QeZM]
// (FP32).  Padding slots past nslot hold r = v = 0 and have no cell
// (E_cr = 0), so their acceleration is exactly 0 and their unguarded stores
// keep them at 0.  NA = kMaxAtom is the generic instantiation (runtime atom
// count, operand table read from global memory, one wavefront per group).
// LDS row pitches (floats): B operands are read as 4 rows x 16 consecutive
// slots per half-wavefront, conflict-free for a pitch of 16 mod 32; the
// accumulator is written as rows 4 apart, conflict-free for 4 x pitch of
// 16 mod 32.
constexpr int kRowB = 80, kRowA = 68;

// Per-wavefront LDS floats for N atoms: the r^n rows, later reused for the
// acceleration rows.
__host__ __device__ constexpr int mfma_wave_lds(int na) {
    return (3 * na + 3) / 4 * 4 * kRowB > (3 * na + 15) / 16 * 16 * kRowA ? (3 * na + 3) / 4 * 4 * kRowB
                                                                         : (3 * na + 15) / 16 * 16 * kRowA;
}

// Compensated FP32 state (IonLayout w = 4): `base` is the tile's first word;
// group j holds the lanes' FP32 words at word 256 j + 4 lane, compensation
// row b (after the ng groups) the lane's 4-bit codes of groups 8b .. 8b+7 at
// words 256 (ng + b) + 4 lane .. +3, two groups per dword.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ f32x4 ldquad(const uint32_t *base, int lane, int j) {
    return __builtin_nontemporal_load(reinterpret_cast<const f32x4 *>(base + (int64_t)j * kPkWords + 4 * lane));
}
__device__ __forceinline__ u32x4 ldcomp(const uint32_t *base, int lane, int ng, int b) {
    return __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(base + (int64_t)(ng + b) * kPkWords + 4 * lane));
}
__device__ __forceinline__ void stquad(f32x4 v, uint32_t *base, int lane, int j) {
    __builtin_nontemporal_store(v, reinterpret_cast<f32x4 *>(base + (int64_t)j * kPkWords + 4 * lane));
}
__device__ __forceinline__ void stcomp(u32x4 v, uint32_t *base, int lane, int ng, int b) {
    __builtin_nontemporal_store(v, reinterpret_cast<u32x4 *>(base + (int64_t)(ng + b) * kPkWords + 4 * lane));
}


template <int NA, int WPB, bool PK>
__device__ __forceinline__ void ion_mfma_block(const SlotArgs<float> &a, const float *__restrict__ aop,
                                               const float *__restrict__ qop, void *__restrict__ rr_,
                                               void *__restrict__ vv_, const int64_t blk) {
    constexpr bool exact = NA != kMaxAtom;
    constexpr int N3X = 3 * NA;
    constexpr int NKX = (N3X + 3) / 4, NAX = (N3X + 15) / 16;
    constexpr int WL = mfma_wave_lds(NA);
    constexpr int W = PK ? 4 : 2;  // dofs per lane vector (IonLayout w)
    constexpr int NVX = (N3X + W - 1) / W;
    constexpr int NBX = (NVX + 7) / 8;  // compensation rows (compensated state)
    // compensated state up to 16 atoms: v is loaded with r, ahead of the MFMA
    // chain (register room; the runtime-N kernel loads it late)
    constexpr bool VPRE = PK && NVX <= 12;
    const int natom = exact ? NA : a.natom;
    const int N3 = 3 * natom;
    const int NK = exact ? NKX : (N3 + 3) / 4, NMA = exact ? NAX : (N3 + 15) / 16;
    const int NV = (N3 + W - 1) / W;
    const int NB = (NV + 7) / 8;
    // compensated state: v's last groups may keep their codes in the free
    // dwords 2 and 3 of r's last compensation row (IonLayout::shared)
    const bool shr = PK && NV % 8 != 0 && NV % 8 <= 4;
    const int NBV = NB - (shr ? 1 : 0);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int grp = lane >> 4, col = lane & 15;
    const int64_t tile = blk * WPB + wave;
    const int64_t ntile = (a.nslot + 63) >> 6;
    extern __shared__ float lds_a[];
    const float *atab = aop;
    if (exact) {
        float *t = lds_a + WPB * WL;
        for (int w = threadIdx.x; w < NAX * NKX * 64; w += 64 * WPB) t[w] = aop[w];
        __syncthreads();
        atab = t;
    }
    if (tile >= ntile) return;  // wave-uniform
    float *buf = lds_a + wave * WL;
    // FP64 state: dof group j of this lane's slot at tb + 128 j; compensated: r's tile words at tw, v's at twv
    const int64_t tb = tile * 64 * W * (int64_t)NV + W * lane;
    const int64_t tw = tile * (int64_t)kPkWords * (NV + NB), twv = tile * (int64_t)kPkWords * (NV + NBV);
    // r^n: FP64 values (FP64 state), or FP32 words + compensation rows
    double r[PK ? 1 : W * NVX];
    float rw[PK ? W * NVX : 1];
    u32x4 rc[PK ? NBX : 1];
    float vw[VPRE ? W * NVX : 1];
    u32x4 vcr[VPRE ? NBX : 1];
    if constexpr (PK) {
#pragma unroll
        for (int b = 0; b < NBX; ++b)
            if (b < NB) rc[b] = ldcomp(static_cast<const uint32_t *>(rr_) + tw, lane, NV, b);
    }
#pragma unroll
    for (int j = 0; j < NVX; ++j)
        if (j < NV) {
            if constexpr (PK) {
                const uint32_t *rb = static_cast<const uint32_t *>(rr_) + tw;
                const f32x4 h4 = ldquad(rb, lane, j);
                for (int h = 0; h < 4; ++h) rw[4 * j + h] = h4[h];
                if constexpr (VPRE) {
                    const uint32_t *vb = static_cast<const uint32_t *>(vv_) + twv;
                    const f32x4 g4 = ldquad(vb, lane, j);
                    for (int h = 0; h < 4; ++h) vw[4 * j + h] = g4[h];
                    if (j % 8 == 0 && j / 8 < NBV) vcr[j / 8] = ldcomp(vb, lane, NV, j / 8);
                }
            } else {
                const f64x2 v2 = ldpair(static_cast<const double *>(rr_) + tb, j);
                r[2 * j] = v2.x;
                r[2 * j + 1] = v2.y;
            }
        }
    const int64_t slot = tile * 64 + lane;
    const bool live = slot < a.nslot;
    size_t cell = 0;
    float Dl[3] = {0.f, 0.f, 0.f}, Eold[3] = {0.f, 0.f, 0.f};
    const SlotE<float> E = slot_e(a);
    if (live) {
        cell = a.cells[slot];
        Dl[0] = a.Dx[cell];
        Dl[1] = a.Dy[cell];
        Dl[2] = a.Dz[cell];
        if (a.ade) {
            Eold[0] = E.o[0][cell];
            Eold[1] = E.o[1][cell];
            Eold[2] = E.o[2][cell];
        }
    }
    // 1. r^n in FP32 as LDS rows (rows 3N .. 4 nk of the last k-tile zero), P_ion.
    //    The FP32 word of compensated state is fl32 of its value.
    float P[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int q = 0; q < N3X; ++q) {
        if (q >= N3) break;
        float rf;
        if constexpr (PK) rf = rw[q];
        else rf = (float)r[q];
        buf[q * kRowB + lane] = rf;
        const int at = q / 3, d = q % 3;
        P[0] += sld(qop, at * 9 + 0 + d) * rf;
        P[1] += sld(qop, at * 9 + 3 + d) * rf;
        P[2] += sld(qop, at * 9 + 6 + d) * rf;
    }
    for (int q = N3; q < 4 * NK; ++q) buf[q * kRowB + lane] = 0.f;
    // 2. E of this lane's slot
    float Ec[3] = {0.f, 0.f, 0.f};
    if (live) {
        float Dc[3];
        rot(a.R, Dl, Dc);
        float Dp[3] = {Dc[0] - P[0], Dc[1] - P[1], Dc[2] - P[2]};
        if (a.ade) {
            float Eo[3], Pn[3];
            rot(a.R, Eold, Eo);
            ade_step(a, slot, Eo, Pn);
            for (int e = 0; e < 3; ++e) Dp[e] -= Pn[e];
        }
        rot(a.A, Dp, Ec);
        float El[3];
        rot_t(a.R, Ec, El);
        E.n[0][cell] = El[0];
        E.n[1][cell] = El[1];
        E.n[2][cell] = El[2];
    }
    __builtin_amdgcn_wave_barrier();
    // 3. -DMM . r^n per 16-slot group on MFMA
    f32x4 acc[4][NAX];
#pragma unroll
    for (int g = 0; g < 4; ++g) {
#pragma unroll
        for (int mt = 0; mt < NAX; ++mt) acc[g][mt] = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kt = 0; kt < NKX; ++kt) {
            if (kt >= NK) break;
            const float b = buf[(4 * kt + grp) * kRowB + 16 * g + col];
#pragma unroll
            for (int mt = 0; mt < NAX; ++mt)
                if (mt < NMA)
                    acc[g][mt] = __builtin_amdgcn_mfma_f32_16x16x4f32(atab[(mt * NK + kt) * 64 + lane], b, acc[g][mt],
                                                                     0, 0, 0);
        }
    }
    __builtin_amdgcn_wave_barrier();
    // 4. acceleration rows back to LDS (dof m, slot 16g + col), then per dof
#pragma unroll
    for (int g = 0; g < 4; ++g)
#pragma unroll
        for (int mt = 0; mt < NAX; ++mt) {
            if (mt >= NMA) break;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int m = 16 * mt + 4 * grp + i;
                if (m < N3) buf[m * kRowA + 16 * g + col] = acc[g][mt][i];
            }
        }
    __builtin_amdgcn_wave_barrier();
    const double dt = a.dt, hdt2 = a.half_dt2;
    const float *qm = qop + 9 * natom;
    // FP64 integration r += v dt + a dt^2/2, v += a dt; dofs past 3N (padding) keep their zeros
    u32x4 vcb;  // compensation row of v holding group j
#pragma unroll
    for (int j = 0; j < NVX; ++j) {
        if (j >= NV) break;
        if constexpr (PK) {
            uint32_t *rb = static_cast<uint32_t *>(rr_) + tw, *vb = static_cast<uint32_t *>(vv_) + twv;
            const bool vin_r = shr && j / 8 == NB - 1;  // v's codes of group j in r's row
            f32x4 v4, r4;
            if constexpr (VPRE) {
                if (j % 8 == 0 && !vin_r) vcb = vcr[j / 8];
                v4 = (f32x4){vw[4 * j], vw[4 * j + 1], vw[4 * j + 2], vw[4 * j + 3]};
            } else {
                if (j % 8 == 0 && !vin_r) vcb = ldcomp(vb, lane, NV, j / 8);
                v4 = ldquad(vb, lane, j);
            }
            // the codes of group j: bits 16 (j % 2) .. +15 of dword (j % 8) / 2
            // (v's in dword 2 + (j % 8) / 2 of r's row when shared)
            const int dw = (j % 8) / 2, sb = 16 * (j % 2);
            uint32_t vc = vin_r ? rc[j / 8][2 + dw] : vcb[dw], rcn = rc[j / 8][dw];
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                const int q = 4 * j + h;
                r4[h] = rw[q];
                if (q >= N3) continue;
                const int at = q / 3, d = q % 3;
                const float af = buf[q * kRowA + lane] + sld(qm, at * 9 + 0 + d) * Ec[0] +
                                 sld(qm, at * 9 + 3 + d) * Ec[1] + sld(qm, at * 9 + 6 + d) * Ec[2];
                const double am = (double)af;
                const double rq = pk_unpack(rw[q], pk_code_of(rcn, sb + 4 * h));
                const double vq = pk_unpack(v4[h], pk_code_of(vc, sb + 4 * h));
                float hw;
                int lo;
                pk_pack(rq + (vq * dt + am * hdt2), &hw, &lo);
                r4[h] = hw;
                rcn = pk_with_code(rcn, sb + 4 * h, lo);
                pk_pack(vq + am * dt, &hw, &lo);
                v4[h] = hw;
                vc = pk_with_code(vc, sb + 4 * h, lo);
            }
            stquad(r4, rb, lane, j);
            stquad(v4, vb, lane, j);
            rc[j / 8][dw] = rcn;
            if (vin_r) rc[j / 8][2 + dw] = vc;
            else vcb[dw] = vc;
            if (j % 8 == 7 || j == NV - 1) {  // row complete (padding codes stored as loaded)
                stcomp(rc[j / 8], rb, lane, NV, j / 8);
                if (!vin_r) stcomp(vcb, vb, lane, NV, j / 8);
            }
        } else {
            double *rd = static_cast<double *>(rr_) + tb, *vd = static_cast<double *>(vv_) + tb;
            f64x2 v2 = ldpair(vd, j);
            f64x2 r2 = {r[2 * j], r[2 * j + 1]};
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int q = 2 * j + h;
                if (q >= N3) break;  // padding dof of an odd 3N stays 0
                const int at = q / 3, d = q % 3;
                const float af = buf[q * kRowA + lane] + sld(qm, at * 9 + 0 + d) * Ec[0] +
                                 sld(qm, at * 9 + 3 + d) * Ec[1] + sld(qm, at * 9 + 6 + d) * Ec[2];
                const double am = (double)af;
                const double vq = v2[h];
                r2[h] = r[q] + (vq * dt + am * hdt2);
                v2[h] = vq + am * dt;
            }
            stpair(r2, rd, j);
            stpair(v2, vd, j);
        }
    }
}

template <typename real>
__global__ __launch_bounds__(256) void ade_update(SlotArgs<real> a) {
    ade_slot(a, (int64_t)blockIdx.x * 256 + threadIdx.x);
}

// Single-domain lattice launches take their work-groups in runs of kIonBand
// consecutive blocks per XCD (band_tile): 8.03-8.07 against 8.27-8.34 ms for
// the 16-atom kernel of the 1024x1024x640 case with blocks in launch order
// (alternating contexts, one process).
constexpr int kIonBand = 16;

template <int NA>
__global__ __launch_bounds__(256) void ion_update_f64(SlotArgs<double> a, const double *__restrict__ dmm,
                                                      const double *__restrict__ qezv,
                                                      const double *__restrict__ qezm, double *__restrict__ rr,
                                                      double *__restrict__ vv, int nblk) {
    const int b = band_tile(blockIdx.x, nblk, kIonBand);
    if (b < 0) return;  // whole work-group
    ion_f64_slot<NA>(a, dmm, qezv, qezm, rr, vv, (int64_t)b * 256 + threadIdx.x);
}

template <int NA, bool PK>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void ion_update_mfma(
    SlotArgs<float> a, const float *__restrict__ aop, const float *__restrict__ qop, void *__restrict__ rr,
    void *__restrict__ vv, int nblk) {
    const int b = band_tile(blockIdx.x, nblk, kIonBand);
    if (b < 0) return;  // whole work-group
    ion_mfma_block<NA, NA == kMaxAtom ? 1 : 4, PK>(a, aop, qop, rr, vv, b);
}

// Work-groups launched for nb blocks: whole runs of kIonBand on all 8 XCDs.
static unsigned ion_grid(int64_t nb) { return (unsigned)((nb + 8 * kIonBand - 1) / (8 * kIonBand) * (8 * kIonBand)); }

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
    ion_f64_slot<NA>(a, a.dmm, a.qezv, a.qezm, static_cast<double *>(a.r), static_cast<double *>(a.v),
                     (int64_t)(blockIdx.x - sld(b.start, rid)) * 256 + threadIdx.x);
}

template <int NA, bool PK>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void ion_update_mfma_batch(
    BatchLaunch<SlotArgs<float>> b) {
    const int rid = (int)sld(b.map, blockIdx.x);
    const SlotArgs<float> a = batch_args(b, rid);
    ion_mfma_block<NA, NA == kMaxAtom ? 1 : 4, PK>(a, a.aop, a.qop, a.r, a.v, blockIdx.x - sld(b.start, rid));
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
    const int nb = (int)((a.nslot + 255) / 256);
    if (!nb) return;
    const dim3 g(ion_grid(nb));
    double *r = static_cast<double *>(a.r), *v = static_cast<double *>(a.v);
#define PHPFDTD_ION_CASE(NN)                                                                              \
    case NN:                                                                                              \
        hipLaunchKernelGGL(ion_update_f64<NN>, g, dim3(256), 0, st, a, a.dmm, a.qezv, a.qezm, r, v, nb);  \
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
            hipLaunchKernelGGL(ion_update_f64<0>, g, dim3(256), 0, st, a, a.dmm, a.qezv, a.qezm, r, v, nb);
    }
#undef PHPFDTD_ION_CASE
}

// Dynamic LDS of the precision-form lattice kernel: per-wavefront transpose
// rows, plus the operand table staged per work-group (exact instantiations).
static size_t mfma_lds_bytes(int na) {
    if (na == kMaxAtom) return (size_t)mfma_wave_lds(na) * 4;
    return ((size_t)4 * mfma_wave_lds(na) + (size_t)mfma_nma(na) * mfma_nk(na) * 64) * 4;
}

// Threads per work-group of the precision-form lattice kernel.
static int mfma_threads(int na) { return na == kMaxAtom ? 64 : 256; }

template <>
void launch_ion<float>(const SlotArgs<float> &a, hipStream_t st) {
    const int64_t ntile = (a.nslot + 63) / 64;
    if (!ntile) return;
    const bool pk = a.pk != 0;  // compensated FP32 state
    // the host lays out aop/qop for the instantiation mfma_inst(natom)
    switch (mfma_inst(a.natom)) {
#define PHPFDTD_MFMA_CASE(NN)                                                                               \
    case NN: {                                                                                              \
        const int wpb = mfma_threads(NN) / 64;                                                              \
        const int nb = (int)((ntile + wpb - 1) / wpb);                                                      \
        const dim3 g(ion_grid(nb)), blk(mfma_threads(NN));                                                  \
        if (pk) hipLaunchKernelGGL((ion_update_mfma<NN, true>), g, blk, mfma_lds_bytes(NN), st, a, a.aop, a.qop, a.r, a.v, nb); \
        else hipLaunchKernelGGL((ion_update_mfma<NN, false>), g, blk, mfma_lds_bytes(NN), st, a, a.aop, a.qop, a.r, a.v, nb); \
        return;                                                                                             \
    }
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
    if (a.natom > 0 && sizeof(real) == 4) {
        const int wpb = mfma_threads(mfma_inst(a.natom)) / 64;
        return ((a.nslot + 63) / 64 + wpb - 1) / wpb;
    }
    return (a.nslot + 255) / 256;
}

// Instantiation key of a domain's slot kernel: 0 for ADE-only materials,
// else the natom template the kernel is compiled for.
template <typename real>
int slot_kind(const SlotArgs<real> &a) {
    if (a.natom == 0) return 0;
    if (sizeof(real) == 4) return mfma_inst(a.natom) + (a.pk ? kPackedKind : 0);
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

// kind: slot_kind; the lattice kernels of a compensated-state batch are the
// PK instantiations (kind + kPackedKind).
template <>
void launch_slot_batch<float>(const BatchLaunch<SlotArgs<float>> &b, int kind, hipStream_t st) {
    if (b.total <= 0) return;
    const dim3 g(b.total), blk(256);
    switch (kind) {
        case 0: hipLaunchKernelGGL(ade_update_batch<float>, g, blk, 0, st, b); return;
#define PHPFDTD_MFMA_CASE(NN)                                                                                      \
    case NN:                                                                                                       \
        hipLaunchKernelGGL((ion_update_mfma_batch<NN, false>), g, dim3(mfma_threads(NN)), mfma_lds_bytes(NN), st, b); \
        return;                                                                                                    \
    case NN + kPackedKind:                                                                                         \
        hipLaunchKernelGGL((ion_update_mfma_batch<NN, true>), g, dim3(mfma_threads(NN)), mfma_lds_bytes(NN), st, b);  \
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

template void launch_ade<double>(const SlotArgs<double> &, hipStream_t);
template void launch_ade<float>(const SlotArgs<float> &, hipStream_t);
template int64_t slot_blocks<double>(const SlotArgs<double> &);
template int64_t slot_blocks<float>(const SlotArgs<float> &);
template int slot_kind<double>(const SlotArgs<double> &);
template int slot_kind<float>(const SlotArgs<float> &);

}  // namespace phpfdtd
