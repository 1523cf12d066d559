// The mixed face kernel with the basis contractions on the matrix cores
// (faces_batch_mfma); compiled as its own translation unit.
#include "face_device.hpp"

namespace umcp {

// ---------------------------------------------------------------------------
// Mixed path with the basis contractions on the matrix cores.
//
// The two contractions of a Gauss point, X[d][k] = sum_c T[p][c][d] P[c][k]
// (forward: the five derivative rows Ru, Rv, Ruu, Ruv, Rvv) and G[c][k] +=
// sum_d TB[p][c][d] D[d][k] (gradient), are about 70 % of the VALU work of
// face_group<float>.  Here they run as v_mfma_f32_16x16x4f32 products over 64
// faces and four Gauss points at a time, in layouts chosen so that no data
// moves between the matrix products and the pointwise math:
//
//   operand layouts of v_mfma_f32_16x16x4f32 (wave64, lane l, q = l / 16,
//   j = l % 16): A[16][4] supplies A[j][q], B[4][16] supplies B[q][j], and
//   the 16 x 16 result leaves rows 4q .. 4q + 3 of column j in the lane.
//
//   forward, per 16-face tile t and coordinate k:  rows = (point q, d),
//     M-tile 0 holds d = 0..3, M-tile 1 holds d = 4 in row 4q (rows 4q+1..3
//     zero); K = control points in steps of four; columns = faces.  Lane
//     (q, j) then holds all five rows X[d][k] of point 4b + q of face 16t + j.
//   pointwise: lane (q, j) evaluates that (face, point) pair.
//   gradient: K = the four points of the batch (one step per row d), B =
//     D[d][k] exactly as the lane holds it; rows = control points; the lane
//     accumulates G[c = 4q .. 4q + 3][k] of face 16t + j.
//
// Products accumulate in FP32 in a fixed order, so results are deterministic
// (not bitwise equal to face_group<float>, whose FMA order differs).
typedef float mx4 __attribute__((ext_vector_type(4)));

constexpr int kMxRow = 80;                  // LDS row stride of the staged control points (bank spread)
constexpr int kMxP = 16 * 3 * kMxRow;       // [16 cp][3][kMxRow] floats
constexpr int kMxBase = 3 * kWave * 2;      // [3][64] doubles
constexpr int kMxGv = 13 * 3 * kWave;       // [ncp][3][64] volume seed
constexpr int kMxLds = kMxP + kMxBase + kMxGv;
static_assert(48 * kWave <= kMxP, "cross-wave sums reuse the control-point rows");
static_assert(13 * 3 * kWave * 4 + kRecScratchBytes <= kMxP * 4, "group records reuse the control-point rows");

__device__ __forceinline__ mx4 mfma4(float a, float b, mx4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <int NCP, int NPT, int KIND, int WS, int FEAT>
__device__ __forceinline__ void face_group_mx(const FaceArgs &a, const double *__restrict__ r, int group, int slot,
                                              float *lds, double acc[4])
{
    constexpr int NKS = (NCP + 3) / 4;   // control-point steps of four
    constexpr int NB = (NPT + 3) / 4;    // batches of four Gauss points
    static_assert(4 * NKS <= 16 && NCP <= 13, "staging holds 16 control-point rows");
    const float *tab = kind_table<KIND, float>();
    const float *tabb = kind_table_b<KIND, float>();
    const float *wts = tab + NPT * 5 * NCP;
    float *sP = lds;                                   // [16][3][kMxRow] relative control points
    double *sBase = (double *)(lds + kMxP);            // [3][64] face origin (control point 0)
    float *sGv = lds + kMxP + kMxBase;                 // [NCP][3][64] volume seed / alpha
    float *sRed = lds;                                 // [48][64] cross-wave sums (after the points)
    const int lane = threadIdx.x & (kWave - 1);
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int q = lane >> 4, j = lane & 15;
    const int nf = a.nf;
    const double al64[3] = {r[3 * a.nv], r[3 * a.nv + 1], r[3 * a.nv + 2]};
    const float al[3] = {(float)al64[0], (float)al64[1], (float)al64[2]};

    // 1. the restraint multipliers, then staging: lane = face, the waves
    // split the control points
    double lam[2];   // dE/dA_tot, dE/dV
    restraint_multipliers<WS * kWave>(a, slot, lam);
    {
        const int f = group * kWave + lane;
        const int fc = f < nf ? f : nf - 1;
        constexpr int kN = (NCP + WS - 1) / WS, kRows = (4 * NKS + WS - 1) / WS;
        double t[kN][3];
        gather_cps<NCP, WS>(a, r, fc, WS == 1 ? 0 : w, t);
        double base[3];
        {
            const int v0 = a.cp[fc];
            for (int k = 0; k < 3; k++) base[k] = r[3 * v0 + k];
            if (a.has_img[fc]) {
                double o[3];
                image_offset(a.img[fc], a.pbc, o);
                for (int k = 0; k < 3; k++) base[k] += o[k];
            }
        }
        if (w == 0)
#pragma unroll
            for (int k = 0; k < 3; k++) sBase[k * kWave + lane] = base[k];
#pragma unroll
        for (int i = 0; i < kRows; i++) {
            const int c = w + i * WS;
            if (c < 4 * NKS)
#pragma unroll
                for (int k = 0; k < 3; k++)
                    sP[(c * 3 + k) * kMxRow + lane] = (i < kN && c < NCP) ? (float)(t[i < kN ? i : 0][k] - base[k]) : 0.f;
        }
    }
    __syncthreads();
    const bool vol = a.rs && a.do_volume;
    if (vol) {
        // volume restraint of this face (see face_group): dE/dV dV/dP of the
        // flat triangle of corner limit points, FP64 on absolute positions,
        // kept in LDS divided by alpha until the write-out; every wave forms
        // the corners, then writes its share of the control points
        const double *W = kind_corner<KIND>();
        double L[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
#pragma unroll
        for (int c = 0; c < NCP; c++)
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const double x = sBase[k * kWave + lane] + (double)sP[(c * 3 + k) * kMxRow + lane];
#pragma unroll
                for (int qq = 0; qq < 3; qq++) L[qq][k] += W[qq * NCP + c] * x;
            }
        const double lamV = lam[1] * (1.0 / 6.0);
        double D[3][3];
#pragma unroll
        for (int qq = 0; qq < 3; qq++) {
            const double *A = L[(qq + 1) % 3], *B = L[(qq + 2) % 3];
            D[qq][0] = lamV * (A[1] * B[2] - A[2] * B[1]) / al64[0];
            D[qq][1] = lamV * (A[2] * B[0] - A[0] * B[2]) / al64[1];
            D[qq][2] = lamV * (A[0] * B[1] - A[1] * B[0]) / al64[2];
        }
        for (int c = w; c < NCP; c += WS)
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const double v = W[c] * D[0][k] + W[NCP + c] * D[1][k] + W[2 * NCP + c] * D[2][k];
                sGv[(c * 3 + k) * kWave + lane] = (float)v;
            }
    }

    // 2. Gauss points, four at a time
    mx4 G[4][3];
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int k = 0; k < 3; k++) G[t][k] = mx4{0.f, 0.f, 0.f, 0.f};
    const double lamA = lam[0];
    const float dudv = (float)a.dudv;
    // per-face inputs of the lane's four faces (16 t + j), loaded up front
    constexpr bool kLip = FEAT != kFeatHelfrich, kC0 = FEAT != kFeatLipid;
    float flw[4], fls[4], flv[4], fc0[4];
    int fcs[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const int f = group * kWave + 16 * t + j;
        fcs[t] = f < nf ? f : nf - 1;
        flw[t] = kLip ? (float)a.lw[fcs[t]] : 0.f;
        fls[t] = kLip ? (float)a.ls[fcs[t]] : 0.f;
        flv[t] = kLip ? (float)a.lv[fcs[t]] : 0.f;
        fc0[t] = kC0 ? (float)a.c0[fcs[t]] : 0.f;
    }
    double e = 0.0;
#pragma unroll 1
    for (int b = w; b < NB; b += WS) {
        // A operands: forward row j = (point j / 4, row j % 4), column q =
        // control point 4s + q; gradient row j = control point, column q = point
        float A0[NKS], A1[NKS], AB[5];
        const int pf = 4 * b + (j >> 2);
#pragma unroll
        for (int s2 = 0; s2 < NKS; s2++) {
            const int c = 4 * s2 + q;
            const bool ok = pf < NPT && c < NCP;
            const int i0 = ok ? (pf * NCP + c) * 5 : 0;
            A0[s2] = ok ? tab[i0 + (j & 3)] : 0.f;
            A1[s2] = (ok && (j & 3) == 0) ? tab[i0 + 4] : 0.f;
        }
        const int p = 4 * b + q;
        const bool pt = p < NPT;
        {
            const bool ok = pt && j < NCP;
            const int i0 = ok ? (p * NCP + j) * 5 : 0;
#pragma unroll
            for (int d = 0; d < 5; d++) AB[d] = ok ? tabb[i0 + d] : 0.f;
        }
        const float wp = pt ? wts[p] : 0.f;
        // per-point inputs of the four faces (clamped point index: the pad
        // point of a 15-point rule reads point 14 and is then skipped)
        const int pc = pt ? p : NPT - 1;
        float pg0[4], pig0[4], prr0[4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            pg0[t] = (float)a.g0[pc * nf + fcs[t]];
            pig0[t] = (float)a.ig0[pc * nf + fcs[t]];
            prr0[t] = FEAT == kFeatGeneral ? (float)a.rurv0[pc * nf + fcs[t]] : 0.f;
        }
        // one 16-face tile at a time: forward products, pointwise math,
        // gradient products (holding all four tiles' products at once costs
        // more in occupancy than it gains in independent work)
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int fl = 16 * t + j;
            mx4 X0[3], X1[3];
#pragma unroll
            for (int k = 0; k < 3; k++) X0[k] = X1[k] = mx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int s2 = 0; s2 < NKS; s2++)
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    const float bv = sP[((4 * s2 + q) * 3 + k) * kMxRow + fl];
                    X0[k] = mfma4(A0[s2], bv, X0[k]);
                    X1[k] = mfma4(A1[s2], bv, X1[k]);
                }
            float D[5][3];
            if (pt) {
                float X[5][3];
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    X[0][k] = X0[k][0] * al[k];
                    X[1][k] = X0[k][1] * al[k];
                    X[2][k] = X0[k][2] * al[k];
                    X[3][k] = X0[k][3] * al[k];
                    X[4][k] = X1[k][0] * al[k];
                }
                const float ee = point_energy_grad<float, FEAT, KIND == 0>(X, wp, pg0[t], pig0[t], prr0[t], flw[t], fls[t],
                                                                flv[t], fc0[t], dudv, lamA, a.ph, D);
                if (group * kWave + fl < nf) e += (double)ee;
            } else {
#pragma unroll
                for (int d = 0; d < 5; d++) D[d][0] = D[d][1] = D[d][2] = 0.f;
            }
#pragma unroll
            for (int d = 0; d < 5; d++)
#pragma unroll
                for (int k = 0; k < 3; k++) G[t][k] = mfma4(AB[d], D[d][k], G[t][k]);
        }
    }

    // 3. energy and dE/dalpha_k = sum_c G[c][k] (r + shift)_ck of this wave's points
    acc[0] += e;
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const int fl = 16 * t + j;
        if (group * kWave + fl >= nf) continue;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            double sg = 0.0, sgp = 0.0;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int c = 4 * q + i;
                if (c < NCP) {
                    sg += (double)G[t][k][i];
                    sgp += (double)G[t][k][i] * (double)sP[(c * 3 + k) * kMxRow + fl];
                }
            }
            acc[1 + k] += sgp + sg * sBase[k * kWave + fl];
        }
    }
    __syncthreads();   // sP is reused for the cross-wave sums below
    if (WS > 1) {
        for (int src = 1; src < WS; src++) {
            if (w == src)
#pragma unroll
                for (int t = 0; t < 4; t++)
#pragma unroll
                    for (int k = 0; k < 3; k++)
#pragma unroll
                        for (int i = 0; i < 4; i++) sRed[((t * 3 + k) * 4 + i) * kWave + lane] = G[t][k][i];
            __syncthreads();
            if (w == 0)
#pragma unroll
                for (int t = 0; t < 4; t++)
#pragma unroll
                    for (int k = 0; k < 3; k++)
#pragma unroll
                        for (int i = 0; i < 4; i++) G[t][k][i] += sRed[((t * 3 + k) * 4 + i) * kWave + lane];
            __syncthreads();
        }
    }
    // the group's per-(face, slot) gradients into LDS over the control-point
    // rows, then its records (one per distinct vertex)
    float *T = sP;
    RecordLists<WS * kWave> rl;
    rl.load(a, slot);
    if (w == 0) {
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int fl = 16 * t + j;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int c = 4 * q + i;
                if (c >= NCP) continue;
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    float v = G[t][k][i];
                    if (vol) v += sGv[(c * 3 + k) * kWave + fl];
                    T[(c * 3 + k) * kWave + fl] = v;
                }
            }
        }
    }
    rl.template sum<float>(a, T, al64, T + 13 * 3 * kWave);
}

// The batched face kernel of the matrix-core mixed path: the same work items,
// partial slots and step counter as faces_batch_kernel.
template <int WS, int OCC, int FEAT, bool PERSIST>
__global__ __launch_bounds__(kWave *WS, OCC) void faces_batch_mfma(BatchFaceArgs B)
{
    __shared__ __align__(16) float lds[kMxLds];
    for (int it = xcd_logical_block(blockIdx.x, gridDim.x); it < B.nitems; it += gridDim.x) {
        const FaceItem item = B.items[it];
        const FaceArgs &a = B.fa[item.fa];
        const double *r = B.r ? B.r : a.r;
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        switch (item.kind) {
        case 0: face_group_mx<12, 15, 0, WS, FEAT>(a, r, item.group, item.slot, lds, acc); break;
        case 1: face_group_mx<11, 16, 1, WS, FEAT>(a, r, item.group, item.slot, lds, acc); break;
        default: face_group_mx<13, 16, 2, WS, FEAT>(a, r, item.group, item.slot, lds, acc); break;
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
        __syncthreads();
    }
}

// point splits 1, 2 and 4 (one, two or four batches of four Gauss points per
// wave); the register-bound flag asks for 2 waves per SIMD
template <int WS, int OCC, int FEAT>
static hipError_t launch_mx_ws(const BatchFaceArgs &a, int nb, bool persist, hipStream_t st)
{
    if (persist)
        hipLaunchKernelGGL((faces_batch_mfma<WS, OCC, FEAT, true>), dim3(nb), dim3(kWave * WS), 0, st, a);
    else
        hipLaunchKernelGGL((faces_batch_mfma<WS, OCC, FEAT, false>), dim3(nb), dim3(kWave * WS), 0, st, a);
    return hipGetLastError();
}

template <int FEAT>
static hipError_t launch_mx(const BatchFaceArgs &a, int nb, bool persist, int ws, int occ2, hipStream_t st)
{
    switch (ws * 2 + (occ2 ? 1 : 0)) {
    case 2: return launch_mx_ws<1, 1, FEAT>(a, nb, persist, st);
    case 3: return launch_mx_ws<1, 2, FEAT>(a, nb, persist, st);
    case 4: return launch_mx_ws<2, 1, FEAT>(a, nb, persist, st);
    case 5: return launch_mx_ws<2, 2, FEAT>(a, nb, persist, st);
    case 8: return launch_mx_ws<4, 1, FEAT>(a, nb, persist, st);
    case 9: return launch_mx_ws<4, 2, FEAT>(a, nb, persist, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_faces_mx(const BatchFaceArgs &a, int nb, bool persist, int ws, int occ2, int feat, hipStream_t st)
{
    switch (feat) {
    case kFeatLipid: return launch_mx<kFeatLipid>(a, nb, persist, ws, occ2, st);
    case kFeatHelfrich: return launch_mx<kFeatHelfrich>(a, nb, persist, ws, occ2, st);
    default: return launch_mx<kFeatGeneral>(a, nb, persist, ws, occ2, st);
    }
}

hipError_t copy_tables_faces_mx(int kind, const double *buf, const float *b32, const double *corner, size_t n,
                                int npt, int ncp)
{
    return copy_tables_here(kind, buf, b32, corner, n, npt, ncp);
}

}  // namespace umcp
