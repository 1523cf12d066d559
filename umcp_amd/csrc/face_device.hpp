// Device code shared by the face kernels' translation units: the Gauss-point
// tables in constant memory, reductions, the per-point energy and its exact
// reverse, and the fused energy + gradient of a 64-face group
// (faces_batch_kernel).  Included by umcp_kernels.hip, umcp_faces64.hip,
// umcp_faces32.hip and umcp_faces_mx.hip, which compile side by side.
//
// Reference computations, alexsodt/UMCP:
//   per-point energy      src/interpLib.C:3641-3881 (fenergy), 3426-3639 (ifenergy)
//   per-point gradient    src/grad.C:246-930 (grad), 6605-7450 (igrad)
#ifndef UMCP_FACE_DEVICE_HPP
#define UMCP_FACE_DEVICE_HPP

#include <hip/hip_runtime.h>

#include <cstdint>

#include "umcp_kernels.hpp"
#include "box_spline.hpp"

namespace umcp {

// The Gauss-point tables depend only on the valence of the face's owner, so
// every surface on the device shares one copy in constant memory (one per
// translation unit: each is its own code object; copy_tables_here fills the
// including unit's copy); wave-
// uniform reads from it become scalar loads whose values feed the FMAs as
// SGPR operands.  Layout per kind: [npt][ncp][5] (the five derivative rows
// Ru, Rv, Ruu, Ruv, Rvv of each control point contiguous, so one scalar load
// brings a control point's coefficients) then [npt] weights.  The gradient
// contraction reads a second copy, so the compiler cannot keep the forward
// pass's coefficients alive in SGPRs across the pointwise math.
constexpr int kTabReg = 15 * 5 * 12, kTabV5 = 16 * 5 * 11, kTabV7 = 16 * 5 * 13;
static __constant__ double c_tab_reg[kTabReg + 15];
static __constant__ double c_tab_v5[kTabV5 + 16];
static __constant__ double c_tab_v7[kTabV7 + 16];
static __constant__ double c_tab_reg_b[kTabReg];
static __constant__ double c_tab_v5_b[kTabV5];
static __constant__ double c_tab_v7_b[kTabV7];
// FP32 copies for the mixed path
static __constant__ float c_tab32_reg[kTabReg + 15];
static __constant__ float c_tab32_v5[kTabV5 + 16];
static __constant__ float c_tab32_v7[kTabV7 + 16];
static __constant__ float c_tab32_reg_b[kTabReg];
static __constant__ float c_tab32_v5_b[kTabV5];
static __constant__ float c_tab32_v7_b[kTabV7];

// position weights of the three corners of a face (the limit points of its
// vertices i, ring[e], ring[e+1]): [3][ncp] per kind
static __constant__ double c_corner_reg[3 * 12];
static __constant__ double c_corner_v5[3 * 11];
static __constant__ double c_corner_v7[3 * 13];

// Fills this translation unit's constant-memory tables of one face kind from
// the reformatted host tables: buf = [npt][ncp][5] then [npt] weights (FP64),
// b32 the same in FP32, corner = [3][ncp].
static inline hipError_t copy_tables_here(int kind, const double *buf, const float *b32, const double *corner,
                                          size_t n, int npt, int ncp)
{
    const void *fwd = kind == 0 ? (const void *)HIP_SYMBOL(c_tab_reg)
                                : (kind == 1 ? (const void *)HIP_SYMBOL(c_tab_v5) : (const void *)HIP_SYMBOL(c_tab_v7));
    const void *bwd = kind == 0 ? (const void *)HIP_SYMBOL(c_tab_reg_b)
                                : (kind == 1 ? (const void *)HIP_SYMBOL(c_tab_v5_b) : (const void *)HIP_SYMBOL(c_tab_v7_b));
    hipError_t e = hipMemcpyToSymbol(fwd, buf, (n + npt) * sizeof(double), 0, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpyToSymbol(bwd, buf, n * sizeof(double), 0, hipMemcpyHostToDevice);
    const void *fwd32 = kind == 0 ? (const void *)HIP_SYMBOL(c_tab32_reg)
                                  : (kind == 1 ? (const void *)HIP_SYMBOL(c_tab32_v5) : (const void *)HIP_SYMBOL(c_tab32_v7));
    const void *bwd32 = kind == 0 ? (const void *)HIP_SYMBOL(c_tab32_reg_b)
                                  : (kind == 1 ? (const void *)HIP_SYMBOL(c_tab32_v5_b) : (const void *)HIP_SYMBOL(c_tab32_v7_b));
    if (e == hipSuccess) e = hipMemcpyToSymbol(fwd32, b32, (n + npt) * sizeof(float), 0, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpyToSymbol(bwd32, b32, n * sizeof(float), 0, hipMemcpyHostToDevice);
    const void *cor = kind == 0 ? (const void *)HIP_SYMBOL(c_corner_reg)
                                : (kind == 1 ? (const void *)HIP_SYMBOL(c_corner_v5) : (const void *)HIP_SYMBOL(c_corner_v7));
    if (e == hipSuccess) e = hipMemcpyToSymbol(cor, corner, 3 * (size_t)ncp * sizeof(double), 0, hipMemcpyHostToDevice);
    return e;
}

template <int KIND>
__device__ __forceinline__ const double *kind_corner()
{
    return KIND == 0 ? c_corner_reg : (KIND == 1 ? c_corner_v5 : c_corner_v7);
}

template <int KIND, typename R>
__device__ __forceinline__ const R *kind_table()
{
    if constexpr (sizeof(R) == 8)
        return (const R *)(KIND == 0 ? c_tab_reg : (KIND == 1 ? c_tab_v5 : c_tab_v7));
    else
        return (const R *)(KIND == 0 ? c_tab32_reg : (KIND == 1 ? c_tab32_v5 : c_tab32_v7));
}
template <int KIND, typename R>
__device__ __forceinline__ const R *kind_table_b()
{
    if constexpr (sizeof(R) == 8)
        return (const R *)(KIND == 0 ? c_tab_reg_b : (KIND == 1 ? c_tab_v5_b : c_tab_v7_b));
    else
        return (const R *)(KIND == 0 ? c_tab32_reg_b : (KIND == 1 ? c_tab32_v5_b : c_tab32_v7_b));
}

// Workgroups are dispatched round-robin over the 8 XCDs; give every XCD a
// contiguous run of logical blocks so neighbouring faces (which share
// vertices) hit the same L2.
__device__ __forceinline__ int xcd_logical_block(int bid, int nb)
{
    constexpr int kXcd = 8;
    int q = nb / kXcd, rem = nb % kXcd;
    int x = bid % kXcd, i = bid / kXcd;
    int start = (x < rem) ? x * (q + 1) : rem * (q + 1) + (x - rem) * q;
    return start + i;
}

// Deterministic reduction of four doubles over a block of kBlock threads;
// thread 0 receives the sums.
template <int kBlock>
__device__ __forceinline__ void block_sum4(double v[4])
{
    __shared__ double red[kBlock / 64][4];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int i = 0; i < 4; i++) v[i] += __shfl_down(v[i], off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int i = 0; i < 4; i++) red[wave][i] = v[i];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int w = 1; w < kBlock / 64; w++)
#pragma unroll
            for (int i = 0; i < 4; i++) v[i] += red[w][i];
}

__device__ __forceinline__ void image_offset(int code, const double pbc[3][3], double out[3])
{
    const int ix = code / 9 - 1, iy = (code / 3) % 3 - 1, iz = code % 3 - 1;
#pragma unroll
    for (int k = 0; k < 3; k++) out[k] = ix * pbc[0][k] + iy * pbc[1][k] + iz * pbc[2][k];
}

// Control points c = c0, c0 + STRIDE, ... of face fc, image-shifted, into
// x[i] (cp c0 + i STRIDE; c0 wave-uniform; entries past NCP - 1 repeat the
// last control point and are the caller's to ignore).  Every index load is
// issued, then every position load, before any of them is used, with no
// branch between them, and the periodic fix-up (faces with has_img, a
// divergent branch) comes last: a wave waits for two dependent memory round
// trips rather than one per control point.
template <int NCP, int STRIDE>
__device__ __forceinline__ void gather_cps(const FaceArgs &a, const double *__restrict__ r, int fc, int c0,
                                           double x[(NCP + STRIDE - 1) / STRIDE][3])
{
    constexpr int kN = (NCP + STRIDE - 1) / STRIDE;
    const int nf = a.nf;
    const bool shifted = a.has_img[fc] != 0;
    int v[kN];
#pragma unroll
    for (int i = 0; i < kN; i++) v[i] = a.cp[min(c0 + i * STRIDE, NCP - 1) * nf + fc];
#pragma unroll
    for (int i = 0; i < kN; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) x[i][k] = r[3 * v[i] + k];
    if (shifted) {
#pragma unroll
        for (int i = 0; i < kN; i++) {
            double o[3];
            image_offset(a.img[min(c0 + i * STRIDE, NCP - 1) * nf + fc], a.pbc, o);
#pragma unroll
            for (int k = 0; k < 3; k++) x[i][k] += o[k];
        }
    }
}

// Control points through the group's distinct vertices: the block stages the
// positions of the group's vertices (at most kStageMax of them) in LDS once,
// then every lane reads its control points through their local indices, so a
// 64-face group makes about 95 scattered position loads instead of 768.  The
// local indices and the vertex list load together, so the positions still
// wait for one round trip only.  A group with more distinct vertices reads
// its control points directly (the choice is uniform over the block).  Every
// thread of the block calls this (it synchronises the block; with NT = 64
// and WAVE, one wave stages its own group and synchronises only itself); sV is
// LDS scratch of kStageMax * 3 doubles.  Output as gather_cps<NCP, STRIDE>.
constexpr int kStageMax = 192;
// LDS written by some lanes of a wave, then read by others of the same wave
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
template <int NCP, int STRIDE, int NT, bool WAVE = false>
__device__ __forceinline__ void stage_cps(const FaceArgs &a, const double *__restrict__ r, int slot, int fc, int c0,
                                          double *sV, double x[(NCP + STRIDE - 1) / STRIDE][3])
{
    static_assert(!WAVE || NT == kWave, "a wave stages with its own 64 lanes");
    constexpr int kN = (NCP + STRIDE - 1) / STRIDE;
    constexpr int kJ = (kStageMax + NT - 1) / NT;
    const int tid = WAVE ? (int)(threadIdx.x & (kWave - 1)) : (int)threadIdx.x;
    const int nf = a.nf;
    const int L0 = a.gl_start[slot];
    const int n = a.gl_start[slot + 1] - L0;
    if (n > kStageMax) {
        if (!WAVE) __syncthreads();
        gather_cps<NCP, STRIDE>(a, r, fc, c0, x);
        return;
    }
    int vj[kJ];
#pragma unroll
    for (int j = 0; j < kJ; j++) {
        const int t = tid + j * NT;
        vj[j] = t < n ? a.lvert[L0 + t] : 0;
    }
    int l[kN];
#pragma unroll
    for (int i = 0; i < kN; i++) l[i] = a.loc[min(c0 + i * STRIDE, NCP - 1) * nf + fc];
    double p[kJ][3];
#pragma unroll
    for (int j = 0; j < kJ; j++)
#pragma unroll
        for (int k = 0; k < 3; k++) p[j][k] = r[3 * vj[j] + k];
#pragma unroll
    for (int j = 0; j < kJ; j++) {
        const int t = tid + j * NT;
        if (t < n)
#pragma unroll
            for (int k = 0; k < 3; k++) sV[3 * t + k] = p[j][k];
    }
    if (WAVE)
        wave_sync();
    else
        __syncthreads();
#pragma unroll
    for (int i = 0; i < kN; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) x[i][k] = sV[3 * l[i] + k];
    if (a.has_img[fc]) {
#pragma unroll
        for (int i = 0; i < kN; i++) {
            double o[3];
            image_offset(a.img[min(c0 + i * STRIDE, NCP - 1) * nf + fc], a.pbc, o);
#pragma unroll
            for (int k = 0; k < 3; k++) x[i][k] += o[k];
        }
    }
}

// The global restraints' energy and multipliers (area: src/interpLib.C:
// 4322-4326; volume: src/minimize.C:90-114) from the per-wave sums of the
// group partials, red[w] = {A, V} of wave w of the finalize block, added in
// wave order: out = {dE/dA_tot, dE/dV, restraint energy, A_tot, V}.
constexpr int kFinalizeThreads = 256;
__device__ __forceinline__ void restraint_form(const FaceArgs &a, const double (*red)[2], double out[5])
{
    #pragma clang fp contract(off)
    double sA = red[0][0], sV = red[0][1];
    for (int q = 1; q < kFinalizeThreads / kWave; q++) {
        sA += red[q][0];
        sV += red[q][1];
    }
    double lamA = 0.0, lamV = 0.0, E = 0.0;
    if (a.do_area && a.KA_g != 0.0) {
        const double st = (sA - a.A0) / a.A0;
        E += a.KA_g * a.A0 * st * st;
        lamA = 2.0 * a.KA_g * st;
    }
    if (a.do_volume && a.KV != 0.0) {
        const double sv = (sV - a.V0) / a.V0;
        E += 0.5 * a.KV * a.V0 * sv * sv;
        lamV = a.KV * sv;
    }
    out[0] = lamA;
    out[1] = lamV;
    out[2] = E;
    out[3] = sA;
    out[4] = sV;
}

// The restraint scalars by one block of kFinalizeThreads threads: thread t
// sums groups t, t + kFinalizeThreads, ... in turn (every load issued
// first), then xor butterflies in each wave and the waves in order, so the
// result does not depend on the launch.  Thread 0 receives out.
__device__ __forceinline__ void restraint_scalars(const FaceArgs &a, double out[5])
{
    #pragma clang fp contract(off)
    constexpr int NT = kFinalizeThreads, kU = 8;
    const double2 *part = reinterpret_cast<const double2 *>(a.rpart);
    double sA = 0.0, sV = 0.0;
    for (int i0 = 0; i0 < a.ngroups; i0 += kU * NT) {
        double2 v[kU];
#pragma unroll
        for (int j = 0; j < kU; j++) v[j] = part[min(i0 + j * NT + (int)threadIdx.x, a.ngroups - 1)];
#pragma unroll
        for (int j = 0; j < kU; j++)
            if (i0 + j * NT + (int)threadIdx.x < a.ngroups) {
                sA += v[j].x;
                sV += v[j].y;
            }
    }
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        sA += __shfl_xor(sA, off, kWave);
        sV += __shfl_xor(sV, off, kWave);
    }
    __shared__ double red[NT / kWave][2];
    if ((threadIdx.x & (kWave - 1)) == 0) {
        red[threadIdx.x / kWave][0] = sA;
        red[threadIdx.x / kWave][1] = sV;
    }
    __syncthreads();
    if (threadIdx.x == 0) restraint_form(a, red, out);
}

// The restraint scalars formed by a face block, in two phases so the
// partials' loads share the block's first memory round trip: load() issues
// every load (call it before the group's staging), finish() sums them in
// the finalize's order and forms the scalars (after the staging).  Finalize
// thread t (of NF) sums groups t, t + NF, ... in turn; face thread i stands
// in for finalize threads i, i + NT, ...  Up to kFoldMaxGroups groups.
template <int NT>
struct RestraintFold {
    static constexpr int NF = kFinalizeThreads, kR = (NF + NT - 1) / NT;
    static constexpr int kG = (kFoldMaxGroups + NF - 1) / NF;
    double2 v[kR][kG];

    __device__ __forceinline__ void load(const FaceArgs &a)
    {
        if (!a.rs || !a.rs_fold) return;   // uniform over the block
        const double2 *part = reinterpret_cast<const double2 *>(a.rpart);
#pragma unroll
        for (int j = 0; j < kR; j++)
#pragma unroll
            for (int u = 0; u < kG; u++) {
                const int t = (int)threadIdx.x + j * NT, i = t + u * NF;
                v[j][u] = part[min(i, a.ngroups - 1)];
            }
    }

    // lam = {dE/dA_tot, dE/dV}: from the finalize's rs, or (a.rs_fold) formed
    // here, and then the slot-0 block also writes rs for the gather and the
    // host.  Every thread of the block calls this.
    __device__ __forceinline__ void finish(const FaceArgs &a, int slot, double lam[2])
    {
        #pragma clang fp contract(off)
        lam[0] = lam[1] = 0.0;
        if (!a.rs) return;
        if (!a.rs_fold) {
            lam[0] = a.rs[0];
            lam[1] = a.rs[1];
            return;
        }
        double sA[kR], sV[kR];
#pragma unroll
        for (int j = 0; j < kR; j++) {
            const int t = (int)threadIdx.x + j * NT;
            sA[j] = sV[j] = 0.0;
#pragma unroll
            for (int u = 0; u < kG; u++)
                if (t < NF && t + u * NF < a.ngroups) {
                    sA[j] += v[j][u].x;
                    sV[j] += v[j][u].y;
                }
        }
        __shared__ double red[NF / kWave][2];
#pragma unroll
        for (int j = 0; j < kR; j++) {
#pragma unroll
            for (int off = 1; off < kWave; off <<= 1) {
                sA[j] += __shfl_xor(sA[j], off, kWave);
                sV[j] += __shfl_xor(sV[j], off, kWave);
            }
            if ((threadIdx.x & (kWave - 1)) == 0 && (int)threadIdx.x + j * NT < NF) {
                red[(threadIdx.x + j * NT) / kWave][0] = sA[j];
                red[(threadIdx.x + j * NT) / kWave][1] = sV[j];
            }
        }
        __syncthreads();
        double o[5];
        restraint_form(a, red, o);
        lam[0] = o[0];
        lam[1] = o[1];
        if (slot == 0 && threadIdx.x == 0)
#pragma unroll
            for (int i = 0; i < 5; i++) a.rs[i] = o[i];
    }
};

// The multipliers in one call (the fold's loads not overlapped with
// anything).  Every thread of the block calls this.
template <int NT>
__device__ __forceinline__ void restraint_multipliers(const FaceArgs &a, int slot, double lam[2])
{
    RestraintFold<NT> rf;
    rf.load(a);
    rf.finish(a, slot, lam);
}

// The gradient records of one 64-face group: T[(c * 3 + k) * 64 + lane] holds
// dE/dP of slot c of the group's face lane (LDS); record L of the group's
// range sums its vertex's incidences in their fixed (slot-major) order in
// FP64 and is scaled by alpha once.  NT threads share the work in two parts:
// RecordLists::load issues the loads of the group's incidence lists (early,
// so their latency overlaps the caller's last LDS work), RecordLists::sum
// stores them in LDS scratch (kRecScratchBytes, disjoint from T), waits for
// the block and forms and writes the records.  T must be complete and
// visible to the block when sum is called.
constexpr int kRecIncMax = 13 * kWave;        // incidences of a group
constexpr int kRecLocMax = 13 * kWave + 1;    // records of a group + 1
constexpr int kRecScratchBytes = kRecIncMax * 2 + kRecLocMax * 4;
template <int NT>
struct RecordLists {
    static constexpr int kI = (kRecIncMax + NT - 1) / NT, kL = (kRecLocMax + NT - 1) / NT;
    uint16_t vi[kI];
    int vl[kL];
    int L0, n, m;

    __device__ __forceinline__ void load(const FaceArgs &a, int slot)
    {
        L0 = a.gl_start[slot];
        n = a.gl_start[slot + 1] - L0;
        const int base = a.linc_start[L0];
        m = a.linc_start[L0 + n] - base;
#pragma unroll
        for (int j = 0; j < kI; j++) {
            const int i = (int)threadIdx.x + j * NT;
            vi[j] = i < m ? a.linc[base + i] : 0;
        }
#pragma unroll
        for (int j = 0; j < kL; j++) {
            const int i = (int)threadIdx.x + j * NT;
            vl[j] = i <= n ? a.linc_start[L0 + i] - base : 0;
        }
    }

    template <typename R>
    __device__ __forceinline__ void sum(const FaceArgs &a, const R *T, const double al[3], void *scratch)
    {
        uint16_t *sInc = (uint16_t *)scratch;
        int *sStart = (int *)((char *)scratch + kRecIncMax * 2);
#pragma unroll
        for (int j = 0; j < kI; j++) {
            const int i = (int)threadIdx.x + j * NT;
            if (i < m) sInc[i] = vi[j];
        }
#pragma unroll
        for (int j = 0; j < kL; j++) {
            const int i = (int)threadIdx.x + j * NT;
            if (i <= n) sStart[i] = vl[j];
        }
        __syncthreads();
        for (int l = (int)threadIdx.x; l < n; l += NT) {
            const int i0 = sStart[l], i1 = sStart[l + 1];
            double s[3] = {0.0, 0.0, 0.0};
            // four incidences in flight, added in list order
            for (int i = i0; i < i1; i += 4) {
                R t[4][3];
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int e = sInc[min(i + q, i1 - 1)];
                    const R *p = T + (e >> 6) * 3 * kWave + (e & (kWave - 1));
#pragma unroll
                    for (int k = 0; k < 3; k++) t[q][k] = p[k * kWave];
                }
#pragma unroll
                for (int q = 0; q < 4; q++)
                    if (i + q < i1)
#pragma unroll
                        for (int k = 0; k < 3; k++) s[k] += (double)t[q][k];
            }
            const size_t L = (size_t)(L0 + l);
            if constexpr (sizeof(R) == 4) {
#pragma unroll
                for (int k = 0; k < 3; k++) a.gbuf32[3 * L + k] = (float)(s[k] * al[k]);
            } else {
#pragma unroll
                for (int k = 0; k < 3; k++) a.gbuf[3 * L + k] = s[k] * al[k];
            }
        }
    }
};

// Reciprocal and square root: IEEE-exact in FP64; the hardware
// single-instruction forms (about 1 ulp) in the FP32 path, whose error budget
// is set by FP32 rounding anyway.
template <typename R>
__device__ __forceinline__ R recip(R x)
{
    if constexpr (sizeof(R) == 4) return __builtin_amdgcn_rcpf(x);
    else return R(1) / x;
}
template <typename R>
__device__ __forceinline__ R root(R x)
{
    if constexpr (sizeof(R) == 4) return __builtin_amdgcn_sqrtf(x);
    else return sqrt(x);
}

// Energy-form feature sets, fixed at compile time so the common cases skip
// unused terms: general (runtime mode, kg and k_reg), lipid mode with kg =
// k_reg = 0 (the reference default), Helfrich mode with kg = k_reg = 0.
constexpr int kFeatGeneral = 0, kFeatLipid = 1, kFeatHelfrich = 2;

// One Gauss point: energy and dE/d(Ru, Rv, Ruu, Ruv, Rvv).
//
// The reference builds the principal curvatures c1,2 = -(a+d -/+ sqrt(SAFETY
// + (a-d)^2 + 4bc))/2 of the shape operator S = I^-1 II and differentiates
// through both; every energy term depends on c1 + c2 = -(a + d) = -tr S and
// c1 c2 = det S - SAFETY/4 only, so this evaluates tr S and det S directly:
//   tr S  = (L E22 - 2 M E12 + N E11) / det I,   det S = (L N - M^2) / det I
// with E = first fundamental form, (L, M, N) = n . (Ruu, Ruv, Rvv) and
// n = (Ru x Rv)/g, g = sqrt(det I) = |Ru x Rv|; the gradient is the exact
// reverse of this chain.  ig0 = 1/g0 of this point.
// REG: the point lies on a regular face.  In the Helfrich form surface::grad
// adds 1/2 kc (c1+c2-c0)^2 dudv w to dE/dg on regular faces
// (src/grad.C:448) although the energy's area element is the fixed g0; igrad
// does not (src/grad.C:6822-6824, under #ifndef FIXED_A).  That term is
// added unless ph.grad_exact asks for the energy's exact derivative.
// The gradient of a point in factored form: dE/dRu, dE/dRv, and the rows
// Ruu, Ruv, Rvv as (dL, dM, dN) times the unit normal n.
template <typename R>
struct PointGrad {
    R du[3], dv[3];
    R dl, dm, dn;
    R n[3];
};

template <typename R, int FEAT, bool REG>
__device__ __forceinline__ R point_energy_grad(const R X[5][3], R w, R g0, R ig0, R rurv0, R lw, R ls, R lv,
                                               R c0f, R dudv, double lamA, const PhysParams &ph, PointGrad<R> &out)
{
    constexpr bool kAnyMode = FEAT == kFeatGeneral;
    constexpr bool kExtras = FEAT == kFeatGeneral;   // kg and k_reg terms
    const R *Ru = X[0], *Rv = X[1], *Suu = X[2], *Suv = X[3], *Svv = X[4];
    const R E11 = Ru[0] * Ru[0] + Ru[1] * Ru[1] + Ru[2] * Ru[2];
    const R E12 = Ru[0] * Rv[0] + Ru[1] * Rv[1] + Ru[2] * Rv[2];
    const R E22 = Rv[0] * Rv[0] + Rv[1] * Rv[1] + Rv[2] * Rv[2];
    const R det = E11 * E22 - E12 * E12;
    const R g = root<R>(det);
    const R idet = recip<R>(det);
    const R ig = g * idet;                       // 1/g
    R Nv[3];
    Nv[0] = Ru[1] * Rv[2] - Ru[2] * Rv[1];
    Nv[1] = Ru[2] * Rv[0] - Ru[0] * Rv[2];
    Nv[2] = Ru[0] * Rv[1] - Ru[1] * Rv[0];
    const R n[3] = {Nv[0] * ig, Nv[1] * ig, Nv[2] * ig};
    const R L = n[0] * Suu[0] + n[1] * Suu[1] + n[2] * Suu[2];
    const R M = n[0] * Suv[0] + n[1] * Suv[1] + n[2] * Suv[2];
    const R N = n[0] * Svv[0] + n[1] * Svv[1] + n[2] * Svv[2];
    const R T = L * E22 - R(2) * M * E12 + N * E11;
    const R H = -T * idet;                       // c1 + c2

    const R kc = (R)ph.kc, KA = (R)ph.KA;
    const R g0w = g0 * w;
    R e, dH, hg = R(0);
    const bool lipid = kAnyMode ? (ph.mode == 0) : (FEAT == kFeatLipid);
    if (lipid) {
        // lipid-area weighted leaflets, no dudv, every type (src/interpLib.C:
        // 3818-3838): sum_j 1/4 kc w W_j (H - s_j)^2 = 1/4 kc w (lw (H - ls)^2
        // + lv), the spread lv fixed per face (no cancellation near H = s_j)
        const R q = R(0.25) * kc * w;
        const R hs = H - ls;
        e = q * (lw * hs * hs + lv);
        dH = R(2) * q * lw * hs;
    } else {
        const R hd = H - c0f;
        const R eh = R(0.5) * kc * hd * hd;
        e = eh * dudv * g0w;
        dH = kc * dudv * g0w * hd;
        if (REG && !ph.grad_exact) hg = eh;   // src/grad.C:448
    }
    // per-point area strain (A - A0)/A0 with A = dudv g w, A0 = dudv g0 w;
    // lamA = dE/dA_tot of the global area restraint (A_tot = sum dudv g w),
    // an FP64 multiplier that the FP32 path adds in FP64 (one rounding)
    const R strain = g * ig0 - R(1);
    e += KA * dudv * g0w * strain * strain;
    R dg;
    if constexpr (sizeof(R) == 8)
        dg = (R(2) * KA * strain + lamA + hg) * dudv * w;
    else
        dg = (R)((double)(R(2) * KA * strain + hg) + lamA) * dudv * w;

    R dL = -dH * E22 * idet, dM = R(2) * dH * E12 * idet, dN = -dH * E11 * idet;
    R ddet = dH * T * idet * idet + dg * R(0.5) * ig;
    R dE11 = -dH * N * idet + ddet * E22;
    R dE22 = -dH * L * idet + ddet * E11;
    R dE12 = R(2) * dH * M * idet - R(2) * ddet * E12;
    if constexpr (kExtras) {
        // Gaussian curvature term kg c1 c2 = kg (det S - SAFETY/4)
        const R K = (L * N - M * M) * idet;
        const R dK = (R)ph.kg * dudv * g0w;
        e += dK * (K - R(0.25e-12));
        dL += dK * N * idet;
        dM -= R(2) * dK * M * idet;
        dN += dK * L * idet;
        const R dd = -dK * K * idet;
        dE11 += dd * E22;
        dE22 += dd * E11;
        dE12 -= R(2) * dd * E12;
        // shear regulariser on cos(angle(Ru, Rv)) (src/interpLib.C:3797-3801)
        const R kreg = (R)ph.k_reg;
        const R s1122 = sqrt(E11 * E22);
        const R val = E12 / s1122;
        const R dv_ = val - rurv0;
        e += kreg * dv_ * dv_;
        const R dval = R(2) * kreg * dv_;
        dE11 -= dval * val * R(0.5) / E11;
        dE22 -= dval * val * R(0.5) / E22;
        dE12 += dval / s1122;
    }
    R Q[3], dNv[3];
#pragma unroll
    for (int k = 0; k < 3; k++) Q[k] = dL * Suu[k] + dM * Suv[k] + dN * Svv[k];
    const R qn = Q[0] * n[0] + Q[1] * n[1] + Q[2] * n[2];
#pragma unroll
    for (int k = 0; k < 3; k++) dNv[k] = (Q[k] - qn * n[k]) * ig;
    // dE/dRu += Rv x dN ; dE/dRv += dN x Ru
    out.du[0] = R(2) * dE11 * Ru[0] + dE12 * Rv[0] + (Rv[1] * dNv[2] - Rv[2] * dNv[1]);
    out.du[1] = R(2) * dE11 * Ru[1] + dE12 * Rv[1] + (Rv[2] * dNv[0] - Rv[0] * dNv[2]);
    out.du[2] = R(2) * dE11 * Ru[2] + dE12 * Rv[2] + (Rv[0] * dNv[1] - Rv[1] * dNv[0]);
    out.dv[0] = R(2) * dE22 * Rv[0] + dE12 * Ru[0] + (dNv[1] * Ru[2] - dNv[2] * Ru[1]);
    out.dv[1] = R(2) * dE22 * Rv[1] + dE12 * Ru[1] + (dNv[2] * Ru[0] - dNv[0] * Ru[2]);
    out.dv[2] = R(2) * dE22 * Rv[2] + dE12 * Ru[2] + (dNv[0] * Ru[1] - dNv[1] * Ru[0]);
    out.dl = dL;
    out.dm = dM;
    out.dn = dN;
#pragma unroll
    for (int k = 0; k < 3; k++) out.n[k] = n[k];
    return e;
}

// the same with the five rows expanded: D[d][k] = dE/dX[d][k]
template <typename R, int FEAT, bool REG>
__device__ __forceinline__ R point_energy_grad(const R X[5][3], R w, R g0, R ig0, R rurv0, R lw, R ls, R lv,
                                               R c0f, R dudv, double lamA, const PhysParams &ph, R D[5][3])
{
    PointGrad<R> pg;
    const R e = point_energy_grad<R, FEAT, REG>(X, w, g0, ig0, rurv0, lw, ls, lv, c0f, dudv, lamA, ph, pg);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        D[0][k] = pg.du[k];
        D[1][k] = pg.dv[k];
        D[2][k] = pg.dl * pg.n[k];
        D[3][k] = pg.dm * pg.n[k];
        D[4][k] = pg.dn * pg.n[k];
    }
    return e;
}

// One point's reverse into the control points' gradient G:
// G[c][k] += sum_d TB[c][d] D[d][k], D = dE/d(Ru, Rv, Ruu, Ruv, Rvv).
template <typename R, int NCP>
__device__ __forceinline__ void reverse_point(R G[NCP][3], const R *TB, const PointGrad<R> &pg)
{
    if constexpr (sizeof(R) == 4) {
        // FP32: the second-derivative rows factored, sum_d TB[c][d]
        // D[d][k] = TB0 du_k + TB1 dv_k + (TB2 dL + TB3 dM + TB4 dN) n_k
        // (12 multiply-adds per control point instead of 15; measured 5 %
        // faster on the vesicle, 14 % on the sheet)
#pragma unroll
        for (int c = 0; c < NCP; c++) {
            const R sn = TB[c * 5 + 2] * pg.dl + TB[c * 5 + 3] * pg.dm + TB[c * 5 + 4] * pg.dn;
#pragma unroll
            for (int k = 0; k < 3; k++) G[c][k] += TB[c * 5] * pg.du[k] + TB[c * 5 + 1] * pg.dv[k] + sn * pg.n[k];
        }
    } else {
        // FP64: the five rows expanded (the factored form removes the
        // register spills but measured no faster here)
        R D[5][3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            D[0][k] = pg.du[k];
            D[1][k] = pg.dv[k];
            D[2][k] = pg.dl * pg.n[k];
            D[3][k] = pg.dm * pg.n[k];
            D[4][k] = pg.dn * pg.n[k];
        }
#pragma unroll
        for (int c = 0; c < NCP; c++) {
#pragma unroll
            for (int k = 0; k < 3; k++) {
                R s = G[c][k];
#pragma unroll
                for (int d = 0; d < 5; d++) s += TB[c * 5 + d] * D[d][k];
                G[c][k] = s;
            }
        }
    }
}

// Fused energy + gradient of 64 faces by one workgroup of WS waves.
// Lane l of every wave owns face group*64 + l; wave w evaluates the Gauss
// points p = w, w + WS, ... so the table row of each point is wave-uniform
// (scalar loads feed the FMAs as SGPR operands).  The waves' partial
// gradients are summed in LDS in a fixed order, and each (face, control
// point) gradient is written to its record: records are slot-major per class
// (slot c of face f at c nf + f), so a wave's stores are contiguous; the
// vertex gather finds a vertex's records through the CSR's record index.
//
// R = double: the FP64 path, on absolute coordinates.
// R = float:  the mixed path.  Control points are formed relative to the
// face's first control point in FP64 and rounded to FP32 (the derivative
// rows of every table sum to zero, so the contraction is translation
// invariant); contraction, curvature and chain rule run in FP32; the energy
// and dE/dalpha accumulate in FP64; the gradient leaves as FP32 and is summed
// per vertex in FP64 by the gather.
template <typename R, int NCP, int NPT, int KIND, int WS, bool PLDS, int FEAT>
__device__ __forceinline__ void face_group(const FaceArgs &a, const double *__restrict__ r, int group, int slot,
                                           R *lds, double (*s_corr)[kWave], double acc[4])
{
    constexpr bool kRel = sizeof(R) == 4;
    const R *tab = kind_table<KIND, R>();
    const R *wts = tab + NPT * 5 * NCP;
    const int lane = threadIdx.x & (kWave - 1);
    // wave index made provably wave-uniform so every table address below is
    // too (the compiler then reads the tables with scalar loads)
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int nf = a.nf;
    const int f = group * kWave + lane;
    const bool live = f < nf;
    const int fc = live ? f : nf - 1;
    const double al64[3] = {r[3 * a.nv], r[3 * a.nv + 1], r[3 * a.nv + 2]};
    const R al[3] = {(R)al64[0], (R)al64[1], (R)al64[2]};
    // the restraint partials' loads first (summed after the staging, so
    // their round trip overlaps the staging's), while few registers are live
    RestraintFold<WS * kWave> rf;
    rf.load(a);
    double base[3] = {0.0, 0.0, 0.0};
    R P[PLDS ? 1 : NCP][3];
    if (PLDS) {
        // the staging scratch lies past the control points' rows (in wave 1's
        // gradient buffer, or the record scratch of a one-wave group)
        double t[(NCP + WS - 1) / WS][3];
        stage_cps<NCP, WS, WS * kWave>(a, r, slot, fc, WS == 1 ? 0 : w, (double *)(lds + 13 * 3 * kWave), t);
        if (kRel) {
            const int v0 = a.cp[fc];
            for (int k = 0; k < 3; k++) base[k] = r[3 * v0 + k];
            if (a.has_img[fc]) {
                double o[3];
                image_offset(a.img[fc], a.pbc, o);
                for (int k = 0; k < 3; k++) base[k] += o[k];
            }
        }
#pragma unroll
        for (int i = 0; i < (NCP + WS - 1) / WS; i++) {
            const int c = w + i * WS;
            if (c < NCP)
#pragma unroll
                for (int k = 0; k < 3; k++) lds[(c * 3 + k) * kWave + lane] = (R)(t[i][k] - base[k]);
        }
        __syncthreads();
    } else {
        double t[NCP][3];
        gather_cps<NCP, 1>(a, r, fc, 0, t);
        if (kRel)
            for (int k = 0; k < 3; k++) base[k] = t[0][k];
#pragma unroll
        for (int c = 0; c < NCP; c++)
#pragma unroll
            for (int k = 0; k < 3; k++) P[c][k] = (R)(t[c][k] - base[k]);
    }
#define PV(c, k) (PLDS ? lds[((c) * 3 + (k)) * kWave + lane] : P[PLDS ? 0 : (c)][k])
    double lam[2];   // dE/dA_tot, dE/dV
    rf.finish(a, slot, lam);

    R G[NCP][3];
#pragma unroll
    for (int c = 0; c < NCP; c++) G[c][0] = G[c][1] = G[c][2] = R(0);
    // s_corr[3][64]: the volume seed's share of dE/dalpha (wave 0), in LDS
    const bool seeded = a.rs && a.do_volume;
    if (seeded && w == 0) {
        // Volume restraint, seeded into wave 0's accumulators before the
        // points (few registers are live here): dE/dV dV/dP of this face's
        // flat triangle of corner limit points L0, L1, L2 (V_face = L0 . (L1 x
        // L2) / 6), on absolute positions in FP64 (the faces' terms cancel only
        // about a common origin).  The write-out multiplies G by alpha, so the
        // seed is divided by it; the volume does not depend on alpha, so the
        // seed's share of the alpha gradient is taken off acc here.
        const double *W = kind_corner<KIND>();
        double L[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
#pragma unroll
        for (int c = 0; c < NCP; c++)
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const double x = base[k] + (double)PV(c, k);
#pragma unroll
                for (int q = 0; q < 3; q++) L[q][k] += W[q * NCP + c] * x;
            }
        const double lamV = lam[1] * (1.0 / 6.0);
        double D[3][3];
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const double *A = L[(q + 1) % 3], *B = L[(q + 2) % 3];
            D[q][0] = lamV * (A[1] * B[2] - A[2] * B[1]);
            D[q][1] = lamV * (A[2] * B[0] - A[0] * B[2]);
            D[q][2] = lamV * (A[0] * B[1] - A[1] * B[0]);
        }
        double corr[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int c = 0; c < NCP; c++)
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const double v = W[c] * D[0][k] + W[NCP + c] * D[1][k] + W[2 * NCP + c] * D[2][k];
                G[c][k] = (R)(v / al64[k]);
                corr[k] += (double)G[c][k] * (base[k] + (double)PV(c, k));
            }
        // kept in LDS over the points (registers are what the points need)
#pragma unroll
        for (int k = 0; k < 3; k++) s_corr[k][lane] = corr[k];
    }
    const R lw = (R)a.lw[fc], ls = (R)a.ls[fc], lv = (R)a.lv[fc], c0f = (R)a.c0[fc];
    const double lamA = lam[0];
    double e = 0.0;
    for (int p = w; p < NPT; p += WS) {
        const R *T = tab + p * 5 * NCP;
        R X[5][3];
#pragma unroll
        for (int d = 0; d < 5; d++)
#pragma unroll
            for (int k = 0; k < 3; k++) X[d][k] = R(0);
#pragma unroll
        for (int c = 0; c < NCP; c++) {
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const R x = PV(c, k);
#pragma unroll
                for (int d = 0; d < 5; d++) X[d][k] += T[c * 5 + d] * x;
            }
        }
#pragma unroll
        for (int d = 0; d < 5; d++)
#pragma unroll
            for (int k = 0; k < 3; k++) X[d][k] *= al[k];
        PointGrad<R> pg;
        const R rr0 = FEAT == kFeatGeneral ? (R)a.rurv0[p * nf + fc] : R(0);
        // the mixed path forms 1/g0 from g0 (the same bits as the stored 1/g0
        // once rounded to FP32; one stream fewer, measured 0.9 us faster),
        // the FP64 path reads it (a divide per point measured slower there)
        const double g0d = a.g0[p * nf + fc];
        const R ig0 = sizeof(R) == 4 ? (R)(1.0 / g0d) : (R)a.ig0[p * nf + fc];
        e += (double)point_energy_grad<R, FEAT, KIND == 0>(X, wts[p], (R)g0d, ig0, rr0,
                                                lw, ls, lv, c0f, (R)a.dudv, lamA, a.ph, pg);
        reverse_point<R, NCP>(G, kind_table_b<KIND, R>() + p * 5 * NCP, pg);
    }
    if (live) {
        acc[0] += e;
        // dE/dalpha_k = sum_c dE/dR_ck (r + shift)_ck is linear in G: each
        // wave adds its own points' share (in FP64)
#pragma unroll
        for (int k = 0; k < 3; k++) {
            double sg = 0.0, sgp = 0.0;
#pragma unroll
            for (int c = 0; c < NCP; c++) {
                sg += (double)G[c][k];
                sgp += (double)G[c][k] * (double)PV(c, k);
            }
            acc[1 + k] += sgp + sg * base[k];
            if (seeded && w == 0) acc[1 + k] -= s_corr[k][lane];
        }
    }
#undef PV
    // Every wave puts its G in its own LDS buffer (the control points' buffer
    // is dead once every wave has finished its points); with several waves,
    // wave w sums the buffers of control points c = w, w + WS, ... in wave
    // order (the order of a sequential sum over the waves) into buffer 0.
    // Then the group's records: one per distinct vertex.
    __syncthreads();
#pragma unroll
    for (int c = 0; c < NCP; c++)
#pragma unroll
        for (int k = 0; k < 3; k++) lds[((w * NCP + c) * 3 + k) * kWave + lane] = G[c][k];
    RecordLists<WS * kWave> rl;
    rl.load(a, slot);
    __syncthreads();
    if (WS > 1) {
        for (int c = w; c < NCP; c += WS)
#pragma unroll
            for (int k = 0; k < 3; k++) {
                R s = lds[(c * 3 + k) * kWave + lane];
#pragma unroll
                for (int q = 1; q < WS; q++) s += lds[((q * NCP + c) * 3 + k) * kWave + lane];
                lds[(c * 3 + k) * kWave + lane] = s;
            }
        __syncthreads();
    }
    // record scratch: past buffer 0 (the other waves' buffers are dead; with
    // one wave the kernel's LDS array has room past it)
    rl.template sum<R>(a, lds, al64, lds + 13 * 3 * kWave);
}

// Fused energy + gradient of 64 faces by WS waves that split the Gauss points
// AND the control points (CPS form).  Lane l owns face group*64 + l in every
// wave.  The control points are staged in LDS relative to control point 0 (in
// FP64 for the FP64 path, rounded to FP32 for the mixed one: the derivative
// rows of every table sum to zero, so the contraction is translation
// invariant and point 0's row drops out).  Round j: wave w evaluates Gauss
// point j WS + w and stages its reverse -- dE/dRu, dE/dRv, the normal and
// (dL, dM, dN) -- in LDS; then every wave applies the round's points to its
// own control points c = w, w + WS, ... (table rows wave-uniform: scalar
// loads feed the FMAs):
//   dE/dP_c += TB0 dE/dRu + TB1 dE/dRv + (TB2 dL + TB3 dM + TB4 dN) n.
// A wave keeps the gradient of only its NCP/WS control points in registers,
// so no cross-wave gradient sum is needed and the footprint allows three to
// four waves per SIMD.  The volume seed (FP64, absolute positions) is added
// to each wave's control points at write-out; dE/dalpha takes only the
// Gauss-point terms (the volume does not depend on alpha).
template <typename R, int NCP, int NPT, int KIND, int WS, int FEAT>
__device__ __forceinline__ void face_group_cps(const FaceArgs &a, const double *__restrict__ r, int group, int slot,
                                               R *sP, R *sD, double acc[4])
{
    constexpr int kOwn = (NCP + WS - 1) / WS;    // control points per wave
    constexpr int kRounds = (NPT + WS - 1) / WS;
    constexpr int kDRows = 12;                   // staged reverse of one point
    const R *tab = kind_table<KIND, R>();
    const R *wts = tab + NPT * 5 * NCP;
    const int lane = threadIdx.x & (kWave - 1);
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int nf = a.nf;
    const int f = group * kWave + lane;
    const bool live = f < nf;
    const int fc = live ? f : nf - 1;
    const double al64[3] = {r[3 * a.nv], r[3 * a.nv + 1], r[3 * a.nv + 2]};
    const R al[3] = {(R)al64[0], (R)al64[1], (R)al64[2]};
    double lam[2];   // dE/dA_tot, dE/dV
    restraint_multipliers<WS * kWave>(a, slot, lam);
    // control point 0 (the origin) and this wave's own control points (the
    // staging scratch lies in the points' reverse rows, unused until then)
    double base[3];
    {
        double t[1][3];
        gather_cps<1, 1>(a, r, fc, 0, t);
        for (int k = 0; k < 3; k++) base[k] = t[0][k];
    }
    {
        double t[kOwn][3];
        stage_cps<NCP, WS, WS * kWave>(a, r, slot, fc, w, (double *)(sP + 13 * 3 * kWave), t);
#pragma unroll
        for (int i = 0; i < kOwn; i++) {
            const int c = w + i * WS;
            if (c >= 1 && c < NCP)
#pragma unroll
                for (int k = 0; k < 3; k++) sP[((c - 1) * 3 + k) * kWave + lane] = (R)(t[i][k] - base[k]);
        }
    }
    __syncthreads();
#define PV(c, k) sP[(((c) - 1) * 3 + (k)) * kWave + lane]
    const R lw = (R)a.lw[fc], ls = (R)a.ls[fc], lv = (R)a.lv[fc], c0f = (R)a.c0[fc];
    const double lamA = lam[0];
    R G[kOwn][3];
#pragma unroll
    for (int i = 0; i < kOwn; i++) G[i][0] = G[i][1] = G[i][2] = R(0);
    double e = 0.0;
    for (int j = 0; j < kRounds; j++) {
        const int p = j * WS + w;
        if (p < NPT) {
            const R *T = tab + p * 5 * NCP;
            R X[5][3];
#pragma unroll
            for (int d = 0; d < 5; d++)
#pragma unroll
                for (int k = 0; k < 3; k++) X[d][k] = R(0);
#pragma unroll
            for (int c = 1; c < NCP; c++)
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    const R x = PV(c, k);
#pragma unroll
                    for (int d = 0; d < 5; d++) X[d][k] += T[c * 5 + d] * x;
                }
#pragma unroll
            for (int d = 0; d < 5; d++)
#pragma unroll
                for (int k = 0; k < 3; k++) X[d][k] *= al[k];
            PointGrad<R> pg;
            const R rr0 = FEAT == kFeatGeneral ? (R)a.rurv0[p * nf + fc] : R(0);
            e += (double)point_energy_grad<R, FEAT, KIND == 0>(X, wts[p], (R)a.g0[p * nf + fc],
                                                               (R)a.ig0[p * nf + fc], rr0, lw, ls, lv, c0f,
                                                               (R)a.dudv, lamA, a.ph, pg);
            R *d = sD + w * kDRows * kWave + lane;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                d[k * kWave] = pg.du[k];
                d[(3 + k) * kWave] = pg.dv[k];
                d[(6 + k) * kWave] = pg.n[k];
            }
            d[9 * kWave] = pg.dl;
            d[10 * kWave] = pg.dm;
            d[11 * kWave] = pg.dn;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < WS; q++) {
            const int pq = j * WS + q;
            if (pq < NPT) {
                const R *TB = kind_table_b<KIND, R>() + pq * 5 * NCP;
                const R *d = sD + q * kDRows * kWave + lane;
                R du[3], dv[3], nn[3];
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    du[k] = d[k * kWave];
                    dv[k] = d[(3 + k) * kWave];
                    nn[k] = d[(6 + k) * kWave];
                }
                const R dl = d[9 * kWave], dm = d[10 * kWave], dn = d[11 * kWave];
#pragma unroll
                for (int i = 0; i < kOwn; i++) {
                    const int c = w + i * WS;
                    if (c < NCP) {
                        const R sn = TB[c * 5 + 2] * dl + TB[c * 5 + 3] * dm + TB[c * 5 + 4] * dn;
#pragma unroll
                        for (int k = 0; k < 3; k++)
                            G[i][k] += TB[c * 5] * du[k] + TB[c * 5 + 1] * dv[k] + sn * nn[k];
                    }
                }
            }
        }
        if (j + 1 < kRounds) __syncthreads();
    }
    // dE/dalpha_k = sum_c dE/dR_ck (r + shift)_ck over this wave's control
    // points (FP64), the Gauss-point terms only
    if (live) {
        acc[0] += e;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            double sg = 0.0, sgp = 0.0;
#pragma unroll
            for (int i = 0; i < kOwn; i++) {
                const int c = w + i * WS;
                if (c < NCP) {
                    sg += (double)G[i][k];
                    if (c > 0) sgp += (double)G[i][k] * (double)PV(c, k);
                }
            }
            acc[1 + k] += sgp + sg * base[k];
        }
    }
    // Volume restraint: dE/dV dV/dP of this face's flat triangle of corner
    // limit points L0, L1, L2 (V_face = L0 . (L1 x L2) / 6), on absolute
    // positions in FP64, added to the wave's own control points; the
    // write-out multiplies by alpha, so it enters divided by alpha
    double seed[kOwn][3];
#pragma unroll
    for (int i = 0; i < kOwn; i++) seed[i][0] = seed[i][1] = seed[i][2] = 0.0;
    if (a.rs && a.do_volume) {
        const double *W = kind_corner<KIND>();
        double L[3][3];
#pragma unroll
        for (int q = 0; q < 3; q++)
#pragma unroll
            for (int k = 0; k < 3; k++) L[q][k] = W[q * NCP] * base[k];
#pragma unroll
        for (int c = 1; c < NCP; c++)
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const double x = base[k] + (double)PV(c, k);
#pragma unroll
                for (int q = 0; q < 3; q++) L[q][k] += W[q * NCP + c] * x;
            }
        const double lamV = lam[1] * (1.0 / 6.0);
        double D[3][3];
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const double *A = L[(q + 1) % 3], *B = L[(q + 2) % 3];
            D[q][0] = lamV * (A[1] * B[2] - A[2] * B[1]);
            D[q][1] = lamV * (A[2] * B[0] - A[0] * B[2]);
            D[q][2] = lamV * (A[0] * B[1] - A[1] * B[0]);
        }
#pragma unroll
        for (int i = 0; i < kOwn; i++) {
            const int c = min(w + i * WS, NCP - 1);
#pragma unroll
            for (int k = 0; k < 3; k++)
                seed[i][k] = (W[c] * D[0][k] + W[NCP + c] * D[1][k] + W[2 * NCP + c] * D[2][k]) / al64[k];
        }
    }
#undef PV
    // the group's per-(face, slot) gradients into LDS (over the staged control
    // points, once every wave is done with them), then its records
    RecordLists<WS * kWave> rl;
    rl.load(a, slot);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kOwn; i++) {
        const int c = w + i * WS;
        if (c < NCP)
#pragma unroll
            for (int k = 0; k < 3; k++) sP[(c * 3 + k) * kWave + lane] = (R)((double)G[i][k] + seed[i][k]);
    }
    rl.template sum<R>(a, sP, al64, sP + 13 * 3 * kWave);
}

// Fused face kernel over a table of work items, one item per 64-face group of
// one face class of one surface.  A surface evaluated alone and every member
// of a batched ensemble run this same kernel; an item writes its partial sums
// to the slot its group has in its surface's own launch, so a surface's
// results do not depend on what else shares the launch.  PERSIST = false: one
// workgroup per item; PERSIST = true: a smaller grid strides over the items.
// B.r / B.step_ctr (non-null for a single surface's own launch) override the
// table's positions and step counter.
template <typename R, int WS, bool PLDS, int OCC, int FEAT, bool PERSIST, bool CPS = false>
__global__ __launch_bounds__(kWave *WS, OCC) void faces_batch_kernel(BatchFaceArgs B)
{
    // CPS: 12 relative control points, then one staged point reverse (12
    // rows) per wave; else 13 control points, then one gradient buffer per
    // wave for the cross-wave sum; both hold the group records' scratch past
    // the first 13 control points' rows
    constexpr int kRecR = (kRecScratchBytes + sizeof(R) - 1) / sizeof(R);
    constexpr int kLdsA = CPS ? (12 * 3 + WS * 12) * kWave : (WS > 1 ? WS : 1) * 13 * 3 * kWave;
    constexpr int kLds = kLdsA > 13 * 3 * kWave + kRecR ? kLdsA : 13 * 3 * kWave + kRecR;
    __shared__ R lds[kLds];
    for (int it = xcd_logical_block(blockIdx.x, gridDim.x); it < B.nitems; it += gridDim.x) {
        const FaceItem item = B.items[it];
        const FaceArgs &a = B.fa[item.fa];
        const double *r = B.r ? B.r : a.r;
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        if constexpr (CPS) {
            R *sD = lds + 12 * 3 * kWave;
            switch (item.kind) {
            case 0: face_group_cps<R, 12, 15, 0, WS, FEAT>(a, r, item.group, item.slot, lds, sD, acc); break;
            case 1: face_group_cps<R, 11, 16, 1, WS, FEAT>(a, r, item.group, item.slot, lds, sD, acc); break;
            default: face_group_cps<R, 13, 16, 2, WS, FEAT>(a, r, item.group, item.slot, lds, sD, acc); break;
            }
        } else {
            __shared__ double s_corr[3][kWave];
            switch (item.kind) {
            case 0: face_group<R, 12, 15, 0, WS, PLDS, FEAT>(a, r, item.group, item.slot, lds, s_corr, acc); break;
            case 1: face_group<R, 11, 16, 1, WS, PLDS, FEAT>(a, r, item.group, item.slot, lds, s_corr, acc); break;
            default: face_group<R, 13, 16, 2, WS, PLDS, FEAT>(a, r, item.group, item.slot, lds, s_corr, acc); break;
            }
        }
        block_sum4<kWave * WS>(acc);
        if (threadIdx.x == 0) {
            double *dst = a.partial + 4 * (size_t)item.slot;
#pragma unroll
            for (int i = 0; i < 4; i++) dst[i] = acc[i];
            int *ctr = B.step_ctr ? B.step_ctr : a.step_ctr;
            if (item.slot == 0 && ctr) ctr[0] += 1;
        }
        if (!PERSIST) break;
        __syncthreads();   // LDS and the reduction scratch are reused by the next item
    }
}

}  // namespace umcp

#endif
