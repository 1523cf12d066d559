// HIP kernels for gfx950 (MI355X) besides the face kernels: the
// deterministic vertex gather (which also finishes the global reductions and
// can apply the Brownian update), the metric kernel used by setg0 and area,
// the restraint pre-pass, single-face energies, the effective-mass kernels,
// device noise and the points on the limit surface; and the launchers.  The
// fused face kernels live in umcp_faces64.hip, umcp_faces32.hip and
// umcp_faces_mx.hip (device code shared through face_device.hpp).
//
// Reference computations, alexsodt/UMCP:
//   reference metric      src/interpLib.C:2969-3111 (setg0)
//   area                  src/interpLib.C:4355-4611
//   Brownian step         docs/timescales.tex:59-67, src/bd.C:1035-1036
#include "face_device.hpp"

namespace umcp {

// the face kernels' translation units (umcp_faces64.hip, umcp_faces32.hip,
// umcp_faces_mx.hip)
hipError_t launch_faces64(const BatchFaceArgs &a, int nb, bool persist, int ws, int flags, int feat, hipStream_t st);
hipError_t launch_faces32(const BatchFaceArgs &a, int nb, bool persist, int ws, int flags, int feat, hipStream_t st);
hipError_t launch_faces_mx(const BatchFaceArgs &a, int nb, bool persist, int ws, int occ2, int feat, hipStream_t st);
hipError_t copy_tables_faces64(int kind, const double *buf, const float *b32, const double *corner, size_t n, int npt,
                               int ncp);
hipError_t copy_tables_faces32(int kind, const double *buf, const float *b32, const double *corner, size_t n, int npt,
                               int ncp);
hipError_t copy_tables_faces_mx(int kind, const double *buf, const float *b32, const double *corner, size_t n,
                                int npt, int ncp);

// Per-point metric: mode 0 writes g0 and RuRv0 (no alpha, as setg0 does);
// mode 1 accumulates sum(dudv g w) and sum(dudv g0 w) into block partials.
template <int NCP, int NPT, int KIND>
__global__ __launch_bounds__(kMetricBlock) void face_metric(FaceArgs a, int mode, double *g0_out,
                                                            double *rurv0_out, double *ig0_out,
                                                            double *partial, int partial_base)
{
    const int lb = xcd_logical_block(blockIdx.x, gridDim.x);
    const int f = lb * kMetricBlock + threadIdx.x;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    if (f < a.nf) {
        const int nf = a.nf;
        const double al[3] = {mode ? a.r[3 * a.nv] : 1.0, mode ? a.r[3 * a.nv + 1] : 1.0,
                              mode ? a.r[3 * a.nv + 2] : 1.0};
        double P[NCP][3];
#pragma unroll
        for (int c = 0; c < NCP; c++) {
            const int v = a.cp[c * nf + f];
            double o[3] = {0.0, 0.0, 0.0};
            if (a.has_img[f]) image_offset(a.img[c * nf + f], a.pbc, o);
#pragma unroll
            for (int k = 0; k < 3; k++) P[c][k] = a.r[3 * v + k] + o[k];
        }
        for (int p = 0; p < NPT; p++) {
            const double *T = kind_table<KIND, double>() + p * 5 * NCP;
            const double wp = kind_table<KIND, double>()[NPT * 5 * NCP + p];
            double Ru[3], Rv[3];
#pragma unroll
            for (int k = 0; k < 3; k++) {
                double su = 0.0, sv = 0.0;
#pragma unroll
                for (int c = 0; c < NCP; c++) {
                    su += T[c * 5 + 0] * P[c][k];
                    sv += T[c * 5 + 1] * P[c][k];
                }
                Ru[k] = su * al[k];
                Rv[k] = sv * al[k];
            }
            const double E11 = Ru[0] * Ru[0] + Ru[1] * Ru[1] + Ru[2] * Ru[2];
            const double E12 = Ru[0] * Rv[0] + Ru[1] * Rv[1] + Ru[2] * Rv[2];
            const double E22 = Rv[0] * Rv[0] + Rv[1] * Rv[1] + Rv[2] * Rv[2];
            const double g = sqrt(E11 * E22 - E12 * E12);
            if (mode == 0) {
                g0_out[p * nf + f] = g;
                rurv0_out[p * nf + f] = E12 / sqrt(E11 * E22);
                ig0_out[p * nf + f] = 1.0 / g;
            } else {
                acc[0] += a.dudv * g * wp;
                acc[1] += a.dudv * a.g0[p * nf + f] * wp;
            }
        }
    }
    if (mode == 1) {
        block_sum4<kMetricBlock>(acc);
        if (threadIdx.x == 0) {
            double *dst = partial + 4 * (size_t)(partial_base + lb);
#pragma unroll
            for (int i = 0; i < 4; i++) dst[i] = acc[i];
        }
    }
}

// ---------------------------------------------------------------------------
// Counter-based normal deviates (Philox4x64-10, Salmon, Moraes, Dror & Shaw,
// SC'11 "Parallel random numbers: as easy as 1, 2, 3"; Box-Muller transform).
// The deviates of (seed, step, vertex) do not depend on how many surfaces
// share a launch or on the launch shape.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void philox4x64_10(unsigned long long x[4], unsigned long long k0,
                                              unsigned long long k1)
{
    constexpr unsigned long long kM0 = 0xD2E7470EE14C6C93ULL, kM1 = 0xCA5A826395121157ULL;
    constexpr unsigned long long kW0 = 0x9E3779B97F4A7C15ULL, kW1 = 0xBB67AE8584CAA73BULL;
#pragma unroll
    for (int round = 0; round < 10; round++) {
        const unsigned long long hi0 = __umul64hi(kM0, x[0]), lo0 = kM0 * x[0];
        const unsigned long long hi1 = __umul64hi(kM1, x[2]), lo1 = kM1 * x[2];
        const unsigned long long y0 = hi1 ^ x[1] ^ k0, y2 = hi0 ^ x[3] ^ k1;
        x[0] = y0;
        x[1] = lo1;
        x[2] = y2;
        x[3] = lo0;
        k0 += kW0;
        k1 += kW1;
    }
}

constexpr unsigned long long kNoiseKey1 = 0x554D4350ULL;   // "UMCP"

// three N(0,1) deviates for (seed, step, vertex v)
__device__ __forceinline__ void normal3(unsigned long long seed, int step, int v, double z[3])
{
    // every operation as written (no FMA contraction): the single-surface and
    // batched kernels that inline this must round identically
    #pragma clang fp contract(off)
    unsigned long long x[4] = {(unsigned long long)v, (unsigned long long)(unsigned)step, 0ULL, 0ULL};
    philox4x64_10(x, seed, kNoiseKey1);
    constexpr double k53 = 1.0 / 9007199254740992.0;   // 2^-53
    const double u0 = ((double)(x[0] >> 11) + 0.5) * k53, u1 = ((double)(x[1] >> 11) + 0.5) * k53;
    const double u2 = ((double)(x[2] >> 11) + 0.5) * k53, u3 = ((double)(x[3] >> 11) + 0.5) * k53;
    constexpr double kTwoPi = 6.283185307179586476925286766559;
    const double rho0 = sqrt(-2.0 * log(u0)), rho1 = sqrt(-2.0 * log(u2));
    z[0] = rho0 * cos(kTwoPi * u1);
    z[1] = rho0 * sin(kTwoPi * u1);
    z[2] = rho1 * cos(kTwoPi * u3);
}

__global__ __launch_bounds__(kGatherBlock) void noise_kernel(unsigned long long seed, int step, int n, double *out)
{
    const int v = blockIdx.x * kGatherBlock + threadIdx.x;
    if (v >= n) return;
    double z[3];
    normal3(seed, step, v, z);
    out[3 * (size_t)v] = z[0];
    out[3 * (size_t)v + 1] = z[1];
    out[3 * (size_t)v + 2] = z[2];
}

// ---------------------------------------------------------------------------
// Vertex-owned gather over the face->vertex CSR (no atomics, fixed order).
//   out_mode 0: write g (3 nv) and dE/dalpha into g[3nv..3nv+2]
//   out_mode 1: Brownian update of r, no g
//   out_mode 2: reductions only
// The block that owns vertex block 0 of a surface also reduces that surface's
// face-kernel partials in fixed order (energy, dE/dalpha) and adds the
// restraint energy.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int philox_step(const GatherArgs &a)
{
    return a.step_ctr ? a.step_ctr[0] - 1 : a.step;
}

__device__ __forceinline__ void gather_reduce(const GatherArgs &a)
{
    // every operation as written (no FMA contraction): the single-surface and
    // batched kernels that inline this must round identically
    #pragma clang fp contract(off)
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < a.npartial; i += kGatherBlock)
#pragma unroll
        for (int q = 0; q < 4; q++) acc[q] += a.partial[4 * (size_t)i + q];
    block_sum4<kGatherBlock>(acc);
    if (threadIdx.x == 0) {
        if (a.rs) acc[0] += a.rs[2];
        // the reference stops on a NaN energy (src/interpLib.C:3546-3550):
        // flag it for the host, whatever path the evaluation takes
        if (a.nan_flag && !isfinite(acc[0])) a.nan_flag[0] = 1;
        if (a.E_out) a.E_out[0] = acc[0];
        if (a.energies) {
            // Brownian steps with device noise index their energies by the
            // step counter; everything else by the caller's step
            const bool by_ctr = a.noise == kNoisePhilox && a.out_mode == 1;
            const int step = by_ctr ? philox_step(a) : a.step;
            if (!by_ctr || (step >= 0 && step < a.energies_cap)) a.energies[step] = acc[0];
        }
        if (a.out_mode == 0 && a.g) {
            a.g[3 * a.nv + 0] = acc[1];
            a.g[3 * a.nv + 1] = acc[2];
            a.g[3 * a.nv + 2] = acc[3];
        }
        if (a.sums) {
#pragma unroll
            for (int q = 0; q < 4; q++) a.sums[q] = acc[q];
        }
    }
}

// kGatherLanes lanes per vertex: lane q sums the records of entries s0 + q,
// s0 + q + kGatherLanes, ... of the vertex's CSR range (neighbouring lanes read
// neighbouring entries of vrec; the records themselves, written slot-major by
// the face kernel, are scattered), then a fixed xor butterfly combines the
// lanes (IEEE addition commutes, so every lane of the group ends with the same
// sum).  The entries are read kGatherUnroll at a time, every load issued
// before the first add (clamped to the range; the adds skip what lies beyond
// it), in the same order.
template <typename GB>
__device__ __forceinline__ void vertex_sum(const GatherArgs &a, int v, int q, bool mine, double g[3])
{
    #pragma clang fp contract(off)
    double g0 = 0.0, g1 = 0.0, g2 = 0.0;
    if (mine) {
        const int s0 = a.vstart[v], s1 = a.vstart[v + 1];
        const GB *src = (const GB *)(sizeof(GB) == 4 ? (const void *)a.gbuf32 : (const void *)a.gbuf);
        constexpr int kU = kGatherUnroll;
        for (int e0 = s0 + q; e0 < s1; e0 += kU * kGatherLanes) {
            int rec[kU];
#pragma unroll
            for (int j = 0; j < kU; j++) rec[j] = a.vrec[min(e0 + j * kGatherLanes, s1 - 1)];
            GB x[kU][3];
#pragma unroll
            for (int j = 0; j < kU; j++) {
                const size_t e = (size_t)rec[j];
                x[j][0] = src[3 * e];
                x[j][1] = src[3 * e + 1];
                x[j][2] = src[3 * e + 2];
            }
#pragma unroll
            for (int j = 0; j < kU; j++)
                if (e0 + j * kGatherLanes < s1) {
                    g0 += (double)x[j][0];
                    g1 += (double)x[j][1];
                    g2 += (double)x[j][2];
                }
        }
    }
#pragma unroll
    for (int off = 1; off < kGatherLanes; off <<= 1) {
        g0 += __shfl_xor(g0, off, kWave);
        g1 += __shfl_xor(g1, off, kWave);
        g2 += __shfl_xor(g2, off, kWave);
    }
    g[0] = g0;
    g[1] = g1;
    g[2] = g2;
}

// The Brownian update of one vertex (out_mode 1): r + (-(D dt/kT) g + sqrt(2
// D dt) xi), xi from the noise array or the device stream of (seed, step, v)
__device__ __forceinline__ void bd_position(const GatherArgs &a, int v, const double r[3], const double g[3],
                                            double out[3])
{
    #pragma clang fp contract(off)
    double z[3];
    if (a.noise == kNoisePhilox) {
        normal3(a.seed, philox_step(a), v, z);
    } else {
        const double *xi = a.xi + 3 * (size_t)v;
        z[0] = xi[0];
        z[1] = xi[1];
        z[2] = xi[2];
    }
#pragma unroll
    for (int k = 0; k < 3; k++) out[k] = r[k] + (-a.mobility_dt * g[k] + a.noise_amp * z[k]);
}

template <typename GB>
__device__ __forceinline__ void gather_vertices(const GatherArgs &a, int vblock)
{
    // every operation as written (no FMA contraction): the single-surface and
    // batched kernels that inline this must round identically
    #pragma clang fp contract(off)
    if (a.out_mode == 2) return;
    const int q = threadIdx.x & (kGatherLanes - 1);
    const int v = vblock * kGatherVerts + (int)threadIdx.x / kGatherLanes;
    const bool mine = v < a.nv;
    double gsum[3];
    vertex_sum<GB>(a, v, q, mine, gsum);
    const double g0 = gsum[0], g1 = gsum[1], g2 = gsum[2];
    if (q != 0) return;
    if (!mine) return;
    if (a.out_mode == 0) {
        a.g[3 * v] = g0;
        a.g[3 * v + 1] = g1;
        a.g[3 * v + 2] = g2;
    } else if (a.out_mode == 3) {
        // momenta of an effective-mass step: friction through the inverse
        // mass (Langevin, OLD_LANGEVIN form), half kick, noise, half kick
        // (src/hd.C:1546-1557, 1600-1605, 1650-1693)
        double z[3];
        if (a.noise == kNoisePhilox) {
            normal3(a.seed, philox_step(a), v, z);
        } else {
            const double *xi = a.xi + 3 * (size_t)v;
            z[0] = xi[0];
            z[1] = xi[1];
            z[2] = xi[2];
        }
        const double gv[3] = {g0, g1, g2};
        const double al[3] = {a.r[3 * a.nv], a.r[3 * a.nv + 1], a.r[3 * a.nv + 2]};
        const double Delta = a.effm_akma * a.effm_dt;
        const double sig = sqrt(2 * a.effm_gamma * a.effm_kT * a.effm_akma * a.effm_dt);
        const double scale = -a.effm_akma * a.effm_dt;
        const bool lang = a.effm_mode == kEffmLangevin;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            double next = lang ? a.pp_in[3 * (size_t)v + c] : 0.0;
            if (lang) next += a.mp[3 * (size_t)v + c] * (a.effm_gamma * scale / (al[c] * al[c]));
            next += -gv[c] * a.effm_akma * a.effm_dt / 2;
            if (lang) next += sig * z[c];
            else next += (1.0 / a.effm_gamma) * (sig * z[c]) / Delta;
            next += -gv[c] * a.effm_akma * a.effm_dt / 2;
            a.pp_out[3 * (size_t)v + c] = next;
        }
    } else {
        double *r = a.r + 3 * (size_t)v;
        const double ro[3] = {r[0], r[1], r[2]};
        double rn[3];
        bd_position(a, v, ro, gsum, rn);
        r[0] = rn[0];
        r[1] = rn[1];
        r[2] = rn[2];
    }
}

// ---------------------------------------------------------------------------
// Momentum-space integrators of src/hd.C with the inverse effective mass
// (src/hd.C:1543-1720, 1943-1947; src/parallel.C:1028-1075).  The inverse is
// a dense nv x nv array in HBM (the reference's truncation keeps nearly every
// entry: the inverse of a mesh mass matrix decays only slowly with distance),
// applied to each Cartesian component; one wave per row, lanes striding over
// the row (coalesced) and a fixed xor butterfly, so the products are
// deterministic.  Operation order follows the reference's expressions
// (contraction off).

__device__ __forceinline__ void effm_row(const EffmArgs &a, int v, const double *x, double out[3])
{
    #pragma clang fp contract(off)
    const int lane = threadIdx.x & (kWave - 1);
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    // kU entries per lane in flight (indices clamped to the row, products
    // beyond it skipped), summed in the same order as one at a time
    constexpr int kU = 8;
    if (a.rowptr) {
        const int e1 = a.rowptr[v + 1];
        for (int e0 = a.rowptr[v] + lane; e0 < e1; e0 += kU * kWave) {
            double m[kU];
            int u[kU];
#pragma unroll
            for (int j = 0; j < kU; j++) {
                const int e = min(e0 + j * kWave, e1 - 1);
                m[j] = a.val[e];
                u[j] = a.col[e];
            }
            double xx[kU][3];
#pragma unroll
            for (int j = 0; j < kU; j++)
#pragma unroll
                for (int k = 0; k < 3; k++) xx[j][k] = x[3 * (size_t)u[j] + k];
#pragma unroll
            for (int j = 0; j < kU; j++)
                if (e0 + j * kWave < e1) {
                    s0 += m[j] * xx[j][0];
                    s1 += m[j] * xx[j][1];
                    s2 += m[j] * xx[j][2];
                }
        }
    } else {
        const double *row = a.Mi + (size_t)v * a.nv;
        for (int u0 = lane; u0 < a.nv; u0 += kU * kWave) {
            double m[kU], xx[kU][3];
#pragma unroll
            for (int j = 0; j < kU; j++) {
                const size_t u = (size_t)min(u0 + j * kWave, a.nv - 1);
                m[j] = row[u];
#pragma unroll
                for (int k = 0; k < 3; k++) xx[j][k] = x[3 * u + k];
            }
#pragma unroll
            for (int j = 0; j < kU; j++)
                if (u0 + j * kWave < a.nv) {
                    s0 += m[j] * xx[j][0];
                    s1 += m[j] * xx[j][1];
                    s2 += m[j] * xx[j][2];
                }
        }
    }
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        s0 += __shfl_xor(s0, off, kWave);
        s1 += __shfl_xor(s1, off, kWave);
        s2 += __shfl_xor(s2, off, kWave);
    }
    out[0] = s0;
    out[1] = s1;
    out[2] = s2;
}

// set-up of the dense inverse (one block per row)
__global__ __launch_bounds__(kEffmBlock) void effm_scatter_kernel(int n, const int *rowptr, const int *col,
                                                                  const double *val, double *A)
{
    const int i = blockIdx.x;
    for (int k = rowptr[i] + (int)threadIdx.x; k < rowptr[i + 1]; k += kEffmBlock) A[(size_t)i * n + col[k]] = val[k];
}

// potri leaves the inverse in the upper triangle of the column-major array,
// i.e. at [j][i] with j >= i in row-major terms; copy it to [i][j]
__global__ __launch_bounds__(kEffmBlock) void effm_symmetrize_kernel(int n, double *A)
{
    const int i = blockIdx.x;
    for (int j = i + 1 + (int)threadIdx.x; j < n; j += kEffmBlock) A[(size_t)i * n + j] = A[(size_t)j * n + i];
}

// |Mi_ij| / sqrt(|Mi_ii Mi_jj|) <= rel_tol -> 0 (src/modes.C:1889-1905); the
// diagonal always stays, so reading it while other rows are written is safe
__global__ __launch_bounds__(kEffmBlock) void effm_truncate_kernel(int n, double *A, double rel_tol, int *rowcount)
{
    __shared__ int cnt[kEffmBlock];
    const int i = blockIdx.x;
    const double di = A[(size_t)i * n + i];
    int kept = 0;
    for (int j = (int)threadIdx.x; j < n; j += kEffmBlock) {
        double *x = A + (size_t)i * n + j;
        const double dj = A[(size_t)j * n + j];
        if (fabs(*x) / sqrt(fabs(di * dj)) > rel_tol) kept++;
        else *x = 0.0;
    }
    cnt[threadIdx.x] = kept;
    __syncthreads();
    for (int off = kEffmBlock / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) cnt[threadIdx.x] += cnt[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) rowcount[i] = cnt[0];
}

// one block per row: chunks of kEffmBlock columns, a block prefix count of
// the non-zero entries gives each kept entry its place
__global__ __launch_bounds__(kEffmBlock) void effm_compact_kernel(int n, const double *A, const int *rowptr,
                                                                  uint16_t *col, double *val)
{
    __shared__ int pre[kEffmBlock];
    const int i = blockIdx.x;
    int base = rowptr[i];
    for (int j0 = 0; j0 < n; j0 += kEffmBlock) {
        const int j = j0 + (int)threadIdx.x;
        const double x = j < n ? A[(size_t)i * n + j] : 0.0;
        const int keep = x != 0.0 ? 1 : 0;
        pre[threadIdx.x] = keep;
        __syncthreads();
        // inclusive scan (Hillis-Steele) over the chunk
        for (int off = 1; off < kEffmBlock; off <<= 1) {
            const int add = (int)threadIdx.x >= off ? pre[threadIdx.x - off] : 0;
            __syncthreads();
            pre[threadIdx.x] += add;
            __syncthreads();
        }
        if (keep) {
            const int at = base + pre[threadIdx.x] - 1;
            col[at] = (uint16_t)j;   // n <= 65536
            val[at] = x;
        }
        base += pre[kEffmBlock - 1];
        __syncthreads();
    }
}

hipError_t launch_effm_compact(int n, const double *A, const int *rowptr, uint16_t *col, double *val, hipStream_t st)
{
    if (n > 65536) return hipErrorInvalidValue;   // 16-bit column indices
    hipLaunchKernelGGL(effm_compact_kernel, dim3(n), dim3(kEffmBlock), 0, st, n, A, rowptr, col, val);
    return hipGetLastError();
}

hipError_t launch_effm_scatter(int n, const int *rowptr, const int *col, const double *val, double *A,
                               hipStream_t st)
{
    hipLaunchKernelGGL(effm_scatter_kernel, dim3(n), dim3(kEffmBlock), 0, st, n, rowptr, col, val, A);
    return hipGetLastError();
}

hipError_t launch_effm_symmetrize(int n, double *A, hipStream_t st)
{
    hipLaunchKernelGGL(effm_symmetrize_kernel, dim3(n), dim3(kEffmBlock), 0, st, n, A);
    return hipGetLastError();
}

hipError_t launch_effm_truncate(int n, double *A, double rel_tol, int *rowcount, hipStream_t st)
{
    hipLaunchKernelGGL(effm_truncate_kernel, dim3(n), dim3(kEffmBlock), 0, st, n, A, rel_tol, rowcount);
    return hipGetLastError();
}

// positions: r += (Mi pp) / alpha^2 Delta (src/hd.C:1699-1720); the momenta
// carried to the next step (Langevin) or cleared (Brownian, src/hd.C:1943-1947);
// the product Mi pp kept for the next step's friction (the reference forms
// the same product again from the same momenta, src/hd.C:1553)
__global__ __launch_bounds__(kEffmBlock) void effm_position_kernel(EffmArgs a)
{
    #pragma clang fp contract(off)
    const int v = blockIdx.x * (kEffmBlock / kWave) + (int)threadIdx.x / kWave;
    if (v >= a.nv) return;   // whole waves leave together
    const int lane = threadIdx.x & (kWave - 1);
    double q[3];
    effm_row(a, v, a.pp_next, q);
    if (lane != 0) return;
    const double al[3] = {a.r[3 * a.nv], a.r[3 * a.nv + 1], a.r[3 * a.nv + 2]};
    const bool lang = a.mode == kEffmLangevin;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        a.r[3 * (size_t)v + c] += q[c] * (1.0 / (al[c] * al[c])) * a.akma * a.dt;
        a.pp_keep[3 * (size_t)v + c] = lang ? a.pp_next[3 * (size_t)v + c] : 0.0;
        if (lang) a.mp[3 * (size_t)v + c] = q[c];
    }
}

__global__ __launch_bounds__(kEffmBlock) void effm_product_kernel(EffmArgs a, const double *pp)
{
    #pragma clang fp contract(off)
    const int v = blockIdx.x * (kEffmBlock / kWave) + (int)threadIdx.x / kWave;
    if (v >= a.nv) return;
    const int lane = threadIdx.x & (kWave - 1);
    double q[3];
    effm_row(a, v, pp, q);
    if (lane != 0) return;
#pragma unroll
    for (int c = 0; c < 3; c++) a.mp[3 * (size_t)v + c] = q[c];
}

hipError_t launch_effm_position(const EffmArgs &a, hipStream_t st)
{
    const int nb = (a.nv + kEffmBlock / kWave - 1) / (kEffmBlock / kWave);
    if (nb == 0) return hipSuccess;
    hipLaunchKernelGGL(effm_position_kernel, dim3(nb), dim3(kEffmBlock), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_effm_product(const EffmArgs &a, const double *pp, hipStream_t st)
{
    const int nb = (a.nv + kEffmBlock / kWave - 1) / (kEffmBlock / kWave);
    if (nb == 0) return hipSuccess;
    hipLaunchKernelGGL(effm_product_kernel, dim3(nb), dim3(kEffmBlock), 0, st, a, pp);
    return hipGetLastError();
}

// Chebyshev solve (the scalable effective mass).  Contraction off, as in the
// products above; a row's four lanes take its entries k = q, q + 4, ... in
// column order and are summed by a fixed butterfly.
__global__ __launch_bounds__(kChebBlock) void cheb_init_kernel(ChebArgs a, const double *b, double *d0,
                                                               double inv_theta)
{
    #pragma clang fp contract(off)
    const int p = blockIdx.x * kChebBlock + (int)threadIdx.x;
    if (p >= a.n) return;
    const double di = a.dinv[p];
    const size_t v = (size_t)a.perm[p];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const double bc = b[3 * v + c];
        a.r[3 * (size_t)p + c] = bc;
        a.x[3 * (size_t)p + c] = 0.0;
        d0[3 * (size_t)p + c] = di * bc * inv_theta;
    }
}

__global__ __launch_bounds__(kChebBlock) void cheb_step_kernel(ChebArgs a, const double *__restrict__ d_in,
                                                               double *__restrict__ d_out, double c1, double c2)
{
    #pragma clang fp contract(off)
    const int lane = threadIdx.x & (kWave - 1);
    const int q = lane >> 4;   // which of the row's four lanes
    // each XCD takes a contiguous run of rows, so the neighbours' d it reads
    // (rows nearby in the bandwidth-reducing order) sit in its own L2
    const int b = xcd_logical_block(blockIdx.x, gridDim.x);
    const int p = b * kChebRows + ((int)threadIdx.x >> 6) * 16 + (lane & 15);
    const bool live = p < a.n;
    const size_t n = (size_t)a.n;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    if (live) {
#pragma unroll 4
        for (int k = q; k < a.width; k += kChebLanes) {
            const int j = a.col[k * n + p];
            const double m = a.val[k * n + p];
            s0 += m * d_in[3 * (size_t)j];
            s1 += m * d_in[3 * (size_t)j + 1];
            s2 += m * d_in[3 * (size_t)j + 2];
        }
    }
    // every lane of the wave takes part in the butterfly
    s0 += __shfl_xor(s0, 16, kWave);
    s1 += __shfl_xor(s1, 16, kWave);
    s2 += __shfl_xor(s2, 16, kWave);
    s0 += __shfl_xor(s0, 32, kWave);
    s1 += __shfl_xor(s1, 32, kWave);
    s2 += __shfl_xor(s2, 32, kWave);
    if (!live || q != 0) return;
    const double di = a.dinv[p];
    const double s[3] = {s0, s1, s2};
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const size_t o = 3 * (size_t)p + c;
        const double dk = d_in[o];
        a.x[o] += dk;
        const double rr = a.r[o] - s[c];
        a.r[o] = rr;
        d_out[o] = c1 * dk + c2 * (di * rr);
    }
}

__global__ __launch_bounds__(kChebBlock) void cheb_final_kernel(ChebArgs a, const double *d, double *out)
{
    #pragma clang fp contract(off)
    const int p = blockIdx.x * kChebBlock + (int)threadIdx.x;
    if (p >= a.n) return;
    const size_t v = (size_t)a.perm[p];
#pragma unroll
    for (int c = 0; c < 3; c++) out[3 * v + c] = a.x[3 * (size_t)p + c] + d[3 * (size_t)p + c];
}

__global__ __launch_bounds__(kChebBlock) void effm_position_solved_kernel(EffmArgs a)
{
    #pragma clang fp contract(off)
    const int v = blockIdx.x * kChebBlock + (int)threadIdx.x;
    if (v >= a.nv) return;
    const double al[3] = {a.r[3 * a.nv], a.r[3 * a.nv + 1], a.r[3 * a.nv + 2]};
    const bool lang = a.mode == kEffmLangevin;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        a.r[3 * (size_t)v + c] += a.mp[3 * (size_t)v + c] * (1.0 / (al[c] * al[c])) * a.akma * a.dt;
        a.pp_keep[3 * (size_t)v + c] = lang ? a.pp_next[3 * (size_t)v + c] : 0.0;
    }
}

hipError_t launch_cheb_init(const ChebArgs &a, const double *b, double *d0, double inv_theta, hipStream_t st)
{
    const int nb = (a.n + kChebBlock - 1) / kChebBlock;
    if (nb == 0) return hipSuccess;
    hipLaunchKernelGGL(cheb_init_kernel, dim3(nb), dim3(kChebBlock), 0, st, a, b, d0, inv_theta);
    return hipGetLastError();
}

hipError_t launch_cheb_step(const ChebArgs &a, const double *d_in, double *d_out, double c1, double c2,
                            hipStream_t st)
{
    const int nb = (a.n + kChebRows - 1) / kChebRows;
    if (nb == 0) return hipSuccess;
    hipLaunchKernelGGL(cheb_step_kernel, dim3(nb), dim3(kChebBlock), 0, st, a, d_in, d_out, c1, c2);
    return hipGetLastError();
}

hipError_t launch_cheb_final(const ChebArgs &a, const double *d, double *out, hipStream_t st)
{
    const int nb = (a.n + kChebBlock - 1) / kChebBlock;
    if (nb == 0) return hipSuccess;
    hipLaunchKernelGGL(cheb_final_kernel, dim3(nb), dim3(kChebBlock), 0, st, a, d, out);
    return hipGetLastError();
}

hipError_t launch_effm_position_solved(const EffmArgs &a, hipStream_t st)
{
    const int nb = (a.nv + kChebBlock - 1) / kChebBlock;
    if (nb == 0) return hipSuccess;
    hipLaunchKernelGGL(effm_position_solved_kernel, dim3(nb), dim3(kChebBlock), 0, st, a);
    return hipGetLastError();
}

template <typename GB>
__global__ __launch_bounds__(kGatherBlock) void vertex_gather(GatherArgs a)
{
    if (blockIdx.x == 0) gather_reduce(a);
    gather_vertices<GB>(a, blockIdx.x);
}

template <typename GB>
__global__ __launch_bounds__(kGatherBlock) void vertex_gather_batch(BatchGatherArgs B)
{
    for (int it = blockIdx.x; it < B.nitems; it += gridDim.x) {
        const GatherItem item = B.items[it];
        const GatherArgs &a = B.ga[item.surf];
        if (item.vblock == 0) gather_reduce(a);
        gather_vertices<GB>(a, item.vblock);
        __syncthreads();   // the reduction scratch is reused by the next item
    }
}

// ---------------------------------------------------------------------------
// Global restraints (area: the GLOBAL_AREA form of src/interpLib.C:4322-4326;
// volume: src/minimize.C:90-114 over the limit-point polyhedron of
// src/interpLib.C:997-1098 / src/volume.C:350-430), as a pre-pass over the
// face kernel's items: A_tot = sum dudv g w over every Gauss point (alpha-
// scaled, as the energy's g) and V = sum over faces of L0 . (L1 x L2) / 6 of
// the face's corner limit points, per 64-face group; a finalize launch sums
// each surface's partials in fixed order and writes the multipliers the face
// kernel reads.
// ---------------------------------------------------------------------------
// One 64-face group, lane = face: the block stages the group's distinct
// vertices in LDS, each wave reads all of a face's control points from there
// into registers and takes the Gauss points p = w, w + WS, ...; the last wave
// also forms the corner limit points and the volume term.
// (Single-surface and batched evaluations run this same kernel, so
// contraction choices cannot differ between them.)
// FP64 throughout, for the mixed path too: the multipliers depend on
// A_tot - A0 and V - V0, small differences of the sums.
// The area and volume terms of one face (lane) from its control points x:
// wave w of WS takes Gauss points w, w + WS, ..., the last wave the volume.
template <int NCP, int NPT, int KIND, int WS>
__device__ __forceinline__ void restraint_points(const FaceArgs &a, const double al[3], const double x[NCP][3],
                                                 bool live, double acc[4])
{
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    double A = 0.0, V = 0.0;
    if (a.do_area) {
        const double *tab = kind_table<KIND, double>();
        for (int p = w; p < NPT; p += WS) {
            const double *T = tab + p * 5 * NCP;
            double Ru[3] = {0.0, 0.0, 0.0}, Rv[3] = {0.0, 0.0, 0.0};
#pragma unroll
            for (int c = 0; c < NCP; c++)
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    Ru[k] += T[c * 5 + 0] * x[c][k];
                    Rv[k] += T[c * 5 + 1] * x[c][k];
                }
#pragma unroll
            for (int k = 0; k < 3; k++) {
                Ru[k] *= al[k];
                Rv[k] *= al[k];
            }
            const double E11 = Ru[0] * Ru[0] + Ru[1] * Ru[1] + Ru[2] * Ru[2];
            const double E12 = Ru[0] * Rv[0] + Ru[1] * Rv[1] + Ru[2] * Rv[2];
            const double E22 = Rv[0] * Rv[0] + Rv[1] * Rv[1] + Rv[2] * Rv[2];
            A += a.dudv * sqrt(E11 * E22 - E12 * E12) * tab[NPT * 5 * NCP + p];
        }
    }
    if (a.do_volume && w == WS - 1) {
        const double *W = kind_corner<KIND>();
        double L[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
#pragma unroll
        for (int c = 0; c < NCP; c++)
#pragma unroll
            for (int k = 0; k < 3; k++)
#pragma unroll
                for (int q = 0; q < 3; q++) L[q][k] += W[q * NCP + c] * x[c][k];
        const double cx = L[1][1] * L[2][2] - L[1][2] * L[2][1];
        const double cy = L[1][2] * L[2][0] - L[1][0] * L[2][2];
        const double cz = L[1][0] * L[2][1] - L[1][1] * L[2][0];
        V = (L[0][0] * cx + L[0][1] * cy + L[0][2] * cz) * (1.0 / 6.0);
    }
    if (live) {
        acc[0] += A;
        acc[1] += V;
    }
}

template <int NCP, int NPT, int KIND, int WS>
__device__ __forceinline__ void restraint_group(const FaceArgs &a, const double *__restrict__ r, int group, int slot,
                                                double *sV, double acc[4])
{
    const int lane = threadIdx.x & (kWave - 1);
    const int f = group * kWave + lane;
    const bool live = f < a.nf;
    const int fc = live ? f : a.nf - 1;
    const double al[3] = {r[3 * a.nv], r[3 * a.nv + 1], r[3 * a.nv + 2]};
    double x[NCP][3];
    stage_cps<NCP, 1, WS * kWave>(a, r, slot, fc, 0, sV, x);
    restraint_points<NCP, NPT, KIND, WS>(a, al, x, live, acc);
}

constexpr int kPrepassWaves = 4;   // waves per 64-face group (Gauss points split)
__global__ __launch_bounds__(kWave *kPrepassWaves) void restraint_prepass_kernel(BatchFaceArgs B)
{
    const int it = xcd_logical_block(blockIdx.x, gridDim.x);
    if (it >= B.nitems) return;
    const FaceItem item = B.items[it];
    const FaceArgs &a = B.fa[item.fa];
    if (!a.rs) return;   // a surface without restraints (uniform over the block)
    __shared__ double sV[kStageMax * 3];
    const double *r = B.r ? B.r : a.r;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    switch (item.kind) {
    case 0: restraint_group<12, 15, 0, kPrepassWaves>(a, r, item.group, item.slot, sV, acc); break;
    case 1: restraint_group<11, 16, 1, kPrepassWaves>(a, r, item.group, item.slot, sV, acc); break;
    default: restraint_group<13, 16, 2, kPrepassWaves>(a, r, item.group, item.slot, sV, acc); break;
    }
    block_sum4<kWave * kPrepassWaves>(acc);
    if (threadIdx.x == 0) {
        a.rpart[2 * (size_t)item.slot] = acc[0];
        a.rpart[2 * (size_t)item.slot + 1] = acc[1];
    }
}

// The restraint scalars of every surface with restraints: block b forms them
// for the surface of item fin[b] (its slot-0 item) from the pre-pass partials.
__global__ __launch_bounds__(kFinalizeThreads) void restraint_finalize_kernel(BatchFaceArgs B)
{
    const FaceArgs &a = B.fa[B.items[B.fin[blockIdx.x]].fa];
    if (!a.rs) return;
    double o[5];
    restraint_scalars(a, o);
    if (threadIdx.x == 0)
        for (int i = 0; i < 5; i++) a.rs[i] = o[i];
}

// One face's energy (FP64, every energy form; no global restraint terms, as
// surface::faceEnergy), one thread per listed face.
template <int NCP, int NPT, int KIND>
__device__ double face_energy_one(const FaceArgs &a, const double *__restrict__ r, int fc, double &area,
                                  double &area0, double &dens)
{
    const int nf = a.nf;
    const double al[3] = {r[3 * a.nv], r[3 * a.nv + 1], r[3 * a.nv + 2]};
    double P[NCP][3];
#pragma unroll
    for (int c = 0; c < NCP; c++) {
        const int v = a.cp[c * nf + fc];
        double o[3] = {0.0, 0.0, 0.0};
        if (a.has_img[fc]) image_offset(a.img[c * nf + fc], a.pbc, o);
#pragma unroll
        for (int k = 0; k < 3; k++) P[c][k] = r[3 * v + k] + o[k];
    }
    const double *tab = kind_table<KIND, double>();
    const double *wts = tab + NPT * 5 * NCP;
    double e = 0.0;
    area = area0 = dens = 0.0;
    for (int p = 0; p < NPT; p++) {
        const double *T = tab + p * 5 * NCP;
        double X[5][3];
#pragma unroll
        for (int d = 0; d < 5; d++)
#pragma unroll
            for (int k = 0; k < 3; k++) {
                double s = 0.0;
#pragma unroll
                for (int c = 0; c < NCP; c++) s += T[c * 5 + d] * P[c][k];
                X[d][k] = s * al[k];
            }
        PointGrad<double> pg;
        e += point_energy_grad<double, kFeatGeneral, KIND == 0>(X, wts[p], a.g0[p * nf + fc], a.ig0[p * nf + fc],
                                                     a.rurv0[p * nf + fc], a.lw[fc], a.ls[fc], a.lv[fc], a.c0[fc], a.dudv, 0.0,
                                                     a.ph, pg);
        // surface::area's terms (src/interpLib.C:4480-4560)
        const double E11 = X[0][0] * X[0][0] + X[0][1] * X[0][1] + X[0][2] * X[0][2];
        const double E12 = X[0][0] * X[1][0] + X[0][1] * X[1][1] + X[0][2] * X[1][2];
        const double E22 = X[1][0] * X[1][0] + X[1][1] * X[1][1] + X[1][2] * X[1][2];
        area += a.dudv * sqrt(E11 * E22 - E12 * E12) * wts[p];
        area0 += a.dudv * a.g0[p * nf + fc] * wts[p];
        // the curvature-energy density penergy subtracts (src/interpLib.C:
        // 3850 regular: 1/2 kc (c1 + c2 - c0)^2 in either form; 3612
        // irregular: the point's energy density, kg c1 c2 in the lipid form)
        // over dudv g0 w; c1 + c2 = -(L E22 - 2 M E12 + N E11) / det and
        // c1 c2 = (L N - M^2) / det - SAFETY / 4 as the reference's roots give
        const double det = E11 * E22 - E12 * E12;
        const double nv[3] = {X[0][1] * X[1][2] - X[0][2] * X[1][1], X[0][2] * X[1][0] - X[0][0] * X[1][2],
                              X[0][0] * X[1][1] - X[0][1] * X[1][0]};
        const double iln = 1.0 / sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
        const double L = (nv[0] * X[2][0] + nv[1] * X[2][1] + nv[2] * X[2][2]) * iln;
        const double M = (nv[0] * X[3][0] + nv[1] * X[3][1] + nv[2] * X[3][2]) * iln;
        const double N = (nv[0] * X[4][0] + nv[1] * X[4][1] + nv[2] * X[4][2]) * iln;
        const double H = -(L * E22 - 2.0 * M * E12 + N * E11) / det;
        const double hd = H - a.c0[fc];
        double x = 0.5 * a.ph.kc * hd * hd;
        if (KIND != 0) {
            const double K = (L * N - M * M) / det - 0.25e-12;
            x = a.ph.mode == 0 ? a.ph.kg * K : x + a.ph.kg * K;
        }
        dens += a.dudv * a.g0[p * nf + fc] * wts[p] * x;
    }
    return e;
}

__global__ __launch_bounds__(kMetricBlock) void face_energy_kernel(const FaceArgs *__restrict__ fa,
                                                                   const int *__restrict__ kinds,
                                                                   const double *__restrict__ r,
                                                                   const int2 *__restrict__ items, int n,
                                                                   double *__restrict__ E)
{
    const int i = blockIdx.x * kMetricBlock + threadIdx.x;
    if (i >= n) return;
    const int2 it = items[i];
    const FaceArgs &a = fa[it.x];
    double e, ar, ar0, dn;
    switch (kinds[it.x]) {
    case 0: e = face_energy_one<12, 15, 0>(a, r, it.y, ar, ar0, dn); break;
    case 1: e = face_energy_one<11, 16, 1>(a, r, it.y, ar, ar0, dn); break;
    default: e = face_energy_one<13, 16, 2>(a, r, it.y, ar, ar0, dn); break;
    }
    E[i] = e;
    E[n + i] = ar;
    E[2 * n + i] = ar0;
    E[3 * n + i] = dn;
}

hipError_t launch_face_energies(const FaceArgs *fa, const int *kinds, int nclass, const double *r, const int2 *items,
                                int n, double *E, hipStream_t st)
{
    (void)nclass;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(face_energy_kernel, dim3((n + kMetricBlock - 1) / kMetricBlock), dim3(kMetricBlock), 0, st,
                       fa, kinds, r, items, n, E);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

int kind_of(int ncp, int npt)
{
    if (ncp == 12 && npt == 15) return 0;
    if (ncp == 11 && npt == 16) return 1;
    if (ncp == 13 && npt == 16) return 2;
    return -1;
}

hipError_t upload_tables(int kind, const double *table, const double *weights, const double *corner, int npt,
                         int ncp)
{
    // host layout [npt][6][ncp] (r_w, r_u, r_v, r_uu, r_uv, r_vv) -> device
    // [npt][ncp][5] without the position row, then the weights
    const size_t n = (size_t)npt * 5 * ncp;
    const size_t cap = kind == 0 ? kTabReg : (kind == 1 ? kTabV5 : kTabV7);
    if (n != cap) return hipErrorInvalidValue;
    double *buf = new double[n + npt];
    for (int p = 0; p < npt; p++)
        for (int c = 0; c < ncp; c++)
            for (int d = 0; d < 5; d++) buf[(size_t)(p * ncp + c) * 5 + d] = table[(size_t)(p * 6 + d + 1) * ncp + c];
    for (int p = 0; p < npt; p++) buf[n + p] = weights[p];
    float *b32 = new float[n + npt];
    for (size_t i = 0; i < n + npt; i++) b32[i] = (float)buf[i];
    hipError_t e = copy_tables_here(kind, buf, b32, corner, n, npt, ncp);
    if (e == hipSuccess) e = copy_tables_faces64(kind, buf, b32, corner, n, npt, ncp);
    if (e == hipSuccess) e = copy_tables_faces32(kind, buf, b32, corner, n, npt, ncp);
    if (e == hipSuccess) e = copy_tables_faces_mx(kind, buf, b32, corner, n, npt, ncp);
    delete[] b32;
    delete[] buf;
    return e;
}

int face_groups(int nf) { return (nf + kWave - 1) / kWave; }
int metric_blocks(int nf) { return (nf + kMetricBlock - 1) / kMetricBlock; }

int features_of(const PhysParams &ph)
{
    if (ph.kg != 0.0 || ph.k_reg != 0.0) return kFeatGeneral;
    return ph.mode == 0 ? kFeatLipid : kFeatHelfrich;
}

hipError_t launch_face_metric(const FaceArgs &a, int kind, int mode, double *g0_out, double *rurv0_out,
                              double *ig0_out, double *partial, int partial_base, hipStream_t st)
{
    if (a.nf == 0) return hipSuccess;
    const int nb = metric_blocks(a.nf);
    switch (kind) {
    case 0:
        hipLaunchKernelGGL((face_metric<12, 15, 0>), dim3(nb), dim3(kMetricBlock), 0, st, a, mode, g0_out,
                           rurv0_out, ig0_out, partial, partial_base);
        break;
    case 1:
        hipLaunchKernelGGL((face_metric<11, 16, 1>), dim3(nb), dim3(kMetricBlock), 0, st, a, mode, g0_out,
                           rurv0_out, ig0_out, partial, partial_base);
        break;
    case 2:
        hipLaunchKernelGGL((face_metric<13, 16, 2>), dim3(nb), dim3(kMetricBlock), 0, st, a, mode, g0_out,
                           rurv0_out, ig0_out, partial, partial_base);
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

int gather_blocks(int nv) { return nv > 0 ? (nv + kGatherVerts - 1) / kGatherVerts : 1; }

hipError_t launch_vertex_gather(const GatherArgs &a, hipStream_t st)
{
    const int nb = (a.out_mode == 2) ? 1 : gather_blocks(a.nv);
    if (a.gbuf32)
        hipLaunchKernelGGL(vertex_gather<float>, dim3(nb), dim3(kGatherBlock), 0, st, a);
    else
        hipLaunchKernelGGL(vertex_gather<double>, dim3(nb), dim3(kGatherBlock), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_faces_batch(const BatchFaceArgs &a, int grid, int ws, int flags, bool mixed, int feat,
                              hipStream_t st)
{
    if (a.nitems <= 0) return hipSuccess;
    const bool persist = grid > 0 && grid < a.nitems;
    const int nb = persist ? grid : a.nitems;
    const int occ2 = (flags & kLaunchOcc2) ? 1 : 0;
    if (mixed && (flags & kLaunchMfma)) return launch_faces_mx(a, nb, persist, ws, occ2, feat, st);
    return mixed ? launch_faces32(a, nb, persist, ws, flags, feat, st)
                 : launch_faces64(a, nb, persist, ws, flags, feat, st);
}

hipError_t launch_gather_batch(const BatchGatherArgs &a, int grid, bool mixed, hipStream_t st)
{
    if (a.nitems <= 0) return hipSuccess;
    const int nb = (grid > 0 && grid < a.nitems) ? grid : a.nitems;
    if (mixed)
        hipLaunchKernelGGL(vertex_gather_batch<float>, dim3(nb), dim3(kGatherBlock), 0, st, a);
    else
        hipLaunchKernelGGL(vertex_gather_batch<double>, dim3(nb), dim3(kGatherBlock), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_restraint_prepass(const BatchFaceArgs &a, hipStream_t st)
{
    if (a.nitems <= 0) return hipSuccess;
    hipLaunchKernelGGL(restraint_prepass_kernel, dim3(a.nitems), dim3(kWave * kPrepassWaves), 0, st, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || a.nfin <= 0) return e;
    hipLaunchKernelGGL(restraint_finalize_kernel, dim3(a.nfin), dim3(kFinalizeThreads), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_noise(unsigned long long seed, int step, int n, double *out, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(noise_kernel, dim3((n + kGatherBlock - 1) / kGatherBlock), dim3(kGatherBlock), 0, st, seed, step, n, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// points on the limit surface
// ---------------------------------------------------------------------------

// The 12 box-spline basis functions and their derivatives d = value, u, v,
// uu, uv, vv at (u, v) in slot order: B[d][y] = 1/12 of the sum over the 15
// monomials u^a v^b w^c (c = 4 - a - b, w = 1 - u - v) of the basis's
// integer coefficient times the monomial's derivative d.  The coefficients
// are compile-time constants: the zero terms fold away.
__device__ __forceinline__ void point_basis(double u, double v, double B[6][12])
{
    const double w = 1.0 - u - v;
    double pu[5], pv[5], pw[5];
    pu[0] = pv[0] = pw[0] = 1.0;
#pragma unroll
    for (int i = 1; i < 5; i++) {
        pu[i] = pu[i - 1] * u;
        pv[i] = pv[i - 1] * v;
        pw[i] = pw[i - 1] * w;
    }
#pragma unroll
    for (int d = 0; d < 6; d++)
#pragma unroll
        for (int y = 0; y < 12; y++) B[d][y] = 0.0;
#pragma unroll
    for (int a = 0; a <= 4; a++)
#pragma unroll
        for (int b = 0; a + b <= 4; b++) {
            const int c = 4 - a - b, m = mono_index(a, b);
            auto M = [&](int pa, int pb, int pc) -> double {
                return (pa < 0 || pb < 0 || pc < 0) ? 0.0 : pu[pa < 0 ? 0 : pa] * pv[pb < 0 ? 0 : pb] * pw[pc < 0 ? 0 : pc];
            };
            double md[6];
            md[0] = M(a, b, c);
            md[1] = a * M(a - 1, b, c) - c * M(a, b, c - 1);
            md[2] = b * M(a, b - 1, c) - c * M(a, b, c - 1);
            md[3] = a * (a - 1) * M(a - 2, b, c) - 2.0 * a * c * M(a - 1, b, c - 1) + c * (c - 1) * M(a, b, c - 2);
            md[4] = a * b * M(a - 1, b - 1, c) - a * c * M(a - 1, b, c - 1) - b * c * M(a, b - 1, c - 1) +
                    c * (c - 1) * M(a, b, c - 2);
            md[5] = b * (b - 1) * M(a, b - 2, c) - 2.0 * b * c * M(a, b - 1, c - 1) + c * (c - 1) * M(a, b, c - 2);
#pragma unroll
            for (int y = 0; y < 12; y++) {
                const double cf = kBoxTable.c[kSlotBasis[y] - 1][m];
                if (cf != 0.0) {
#pragma unroll
                    for (int d = 0; d < 6; d++) B[d][y] += cf * md[d];
                }
            }
        }
#pragma unroll
    for (int d = 0; d < 6; d++)
#pragma unroll
        for (int y = 0; y < 12; y++) B[d][y] /= 12.0;
}

// Third derivatives of the 12 box-spline basis functions at (u, v) in slot
// order: B3[d][y], d = uuu, uuv, uvv, vvv (pgrad's (u, v) gradient,
// src/grad.C:3620-3700).  On the barycentric monomials u^a v^b w^c (w = 1 - u
// - v) d/du acts as del_u - del_w and d/dv as del_v - del_w.
__device__ __forceinline__ double falling(int n, int k)
{
    double r = 1.0;
    for (int t = 0; t < k; t++) r *= (double)(n - t);
    return r;
}

__device__ __forceinline__ void point_basis3(double u, double v, double B3[4][12])
{
    const double w = 1.0 - u - v;
    double pu[5], pv[5], pw[5];
    pu[0] = pv[0] = pw[0] = 1.0;
#pragma unroll
    for (int i = 1; i < 5; i++) {
        pu[i] = pu[i - 1] * u;
        pv[i] = pv[i - 1] * v;
        pw[i] = pw[i - 1] * w;
    }
#pragma unroll
    for (int d = 0; d < 4; d++)
#pragma unroll
        for (int y = 0; y < 12; y++) B3[d][y] = 0.0;
    constexpr int kI[4] = {3, 2, 1, 0};   // orders along u (the rest along v)
    constexpr double kBin[4][4] = {{1, 0, 0, 0}, {1, 1, 0, 0}, {1, 2, 1, 0}, {1, 3, 3, 1}};
#pragma unroll
    for (int a = 0; a <= 4; a++)
#pragma unroll
        for (int b = 0; a + b <= 4; b++) {
            const int c = 4 - a - b, m = mono_index(a, b);
            double md[4];
#pragma unroll
            for (int d = 0; d < 4; d++) {
                const int I = kI[d], J = 3 - kI[d];
                double s = 0.0;
#pragma unroll
                for (int k = 0; k <= I; k++)
#pragma unroll
                    for (int l = 0; l <= J; l++) {
                        const int e = (I - k) + (J - l);   // order along w
                        if (k > a || l > b || e > c) continue;
                        const double sg = (e & 1) ? -1.0 : 1.0;
                        s += sg * kBin[I][k] * kBin[J][l] * falling(a, k) * falling(b, l) * falling(c, e) *
                             pu[a - k] * pv[b - l] * pw[c - e];
                    }
                md[d] = s;
            }
#pragma unroll
            for (int y = 0; y < 12; y++) {
                const double cf = kBoxTable.c[kSlotBasis[y] - 1][m];
                if (cf != 0.0) {
#pragma unroll
                    for (int d = 0; d < 4; d++) B3[d][y] += cf * md[d];
                }
            }
        }
#pragma unroll
    for (int d = 0; d < 4; d++)
#pragma unroll
        for (int y = 0; y < 12; y++) B3[d][y] /= 12.0;
}

__device__ __forceinline__ void cross_d(const double a[3], const double b[3], double c[3])
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

// weights of irregular control point x for the first ND derivatives
template <int ND>
__device__ __forceinline__ void irr_weights(const double *__restrict__ map, int ncp, int x, const double B[6][12],
                                            const double sc[6], double t[ND])
{
#pragma unroll
    for (int d = 0; d < ND; d++) t[d] = 0.0;
#pragma unroll
    for (int y = 0; y < 12; y++) {
        const double mw = map[y * ncp + x];
#pragma unroll
        for (int d = 0; d < ND; d++) t[d] += mw * B[d][y];
    }
#pragma unroll
    for (int d = 0; d < ND; d++) t[d] *= sc[d];
}

// One point of evaluateRNRM or pointGradient.  IRR: the control-point weights
// of an irregular face are the sub-patch map applied to the basis, t[d][x] =
// s_d sum_y map[y][x] B[d][y]; on regular faces t[d][x] = B[d][x].
// surface::ru, rv, g and c at one point from its contractions X (unscaled
// positions; derivatives in the face's parameters): the tangents, shape
// operator and principal curvatures of the alpha-scaled surface, and the
// reference's g — |lru x lrv| of unscaled positions, on irregular faces with
// the sub-patch derivatives scaled by the face's own u and v
// (src/uv_map.C:884-1106, 1347-1700, 2361-2990)
__device__ void point_geometry_out(const PointArgs &a, int p, const PointRec &pr, bool irr, const double al[3],
                                   const double X[6][3])
{
    double gu[3], gv[3], gn[3];
    const double mu = irr ? pr.fu / pr.su : 1.0, mv = irr ? pr.fv / pr.sv : 1.0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        gu[k] = X[1][k] * mu;
        gv[k] = X[2][k] * mv;
    }
    cross_d(gu, gv, gn);
    double Ru[3], Rv[3], Suu[3], Suv[3], Svv[3], n[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        Ru[k] = al[k] * X[1][k];
        Rv[k] = al[k] * X[2][k];
        Suu[k] = al[k] * X[3][k];
        Suv[k] = al[k] * X[4][k];
        Svv[k] = al[k] * X[5][k];
    }
    cross_d(Ru, Rv, n);
    const double ln = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (ln > 0.0) {
        n[0] /= ln;
        n[1] /= ln;
        n[2] /= ln;
    }
    const double E = Ru[0] * Ru[0] + Ru[1] * Ru[1] + Ru[2] * Ru[2];
    const double F = Ru[0] * Rv[0] + Ru[1] * Rv[1] + Ru[2] * Rv[2];
    const double G = Rv[0] * Rv[0] + Rv[1] * Rv[1] + Rv[2] * Rv[2];
    const double g2 = E * G - F * F;
    const double L = Suu[0] * n[0] + Suu[1] * n[1] + Suu[2] * n[2];
    const double M = Suv[0] * n[0] + Suv[1] * n[1] + Suv[2] * n[2];
    const double N = Svv[0] * n[0] + Svv[1] * n[1] + Svv[2] * n[2];
    // shape operator I^-1 II in the (u, v) basis
    const double sa = (L * G - M * F) / g2, sb = (M * G - N * F) / g2;
    const double sc = (M * E - L * F) / g2, sd = (N * E - M * F) / g2;
    const double fac = sqrt(1e-12 + sa * sa + 4.0 * sb * sc - 2.0 * sa * sd + sd * sd);   // SAFETY, src/interp.h:14
    const double c1 = -0.5 * (sa + sd - fac), c2 = -0.5 * (sa + sd + fac);
    double v1[2], v2[2];
    if (fabs(sc) > 1e-6) {
        v1[0] = (sa - sd - fac) / (2.0 * sc);
        v2[0] = (sa - sd + fac) / (2.0 * sc);
        v1[1] = v2[1] = 1.0;
    } else if (fabs(sb) > 1e-6) {
        v1[0] = -sb / (sa - sd);
        v1[1] = 1.0;
        v2[0] = 1.0;
        v2[1] = 0.0;
    } else {
        v1[0] = 1.0;
        v1[1] = 0.0;
        v2[0] = -F / sqrt(E) / sqrt(G);
        v2[1] = 1.0;
    }
    const double l1 = sqrt(v1[0] * v1[0] + v1[1] * v1[1]), l2 = sqrt(v2[0] * v2[0] + v2[1] * v2[1]);
    double *o = a.geom + (size_t)p * kPointGeom;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        o[k] = Ru[k];
        o[3 + k] = Rv[k];
    }
    o[6] = sqrt(gn[0] * gn[0] + gn[1] * gn[1] + gn[2] * gn[2]);
    o[7] = c1 + c2;
    o[8] = c1 * c2;
    o[9] = c1;
    o[10] = c2;
    o[11] = v1[0] / l1;
    o[12] = v1[1] / l1;
    o[13] = v2[0] / l2;
    o[14] = v2[1] / l2;
}

// Mean curvature H = c1 + c2 = -(L E22 - 2 M E12 + N E11) / det I of an
// alpha-scaled point (Ru, Rv, Ruu, Ruv, Rvv) and the exact reverse of it for a
// given dE/dH (no area-element term): D[d][k] = dE/dX_d,k, X = Ru, Rv, Ruu,
// Ruv, Rvv.  c1 + c2 of the reference's roots is -(a + d) exactly, whatever
// SAFETY (src/grad.C:381-399), so the trace form is the reference's H.
struct CurvPoint {
    double E11, E12, E22, idet, iln, n[3], L, M, N, T, H;
};

__device__ __forceinline__ CurvPoint curv_setup(const double Ru[3], const double Rv[3], const double Suu[3],
                                               const double Suv[3], const double Svv[3])
{
    CurvPoint c;
    c.E11 = Ru[0] * Ru[0] + Ru[1] * Ru[1] + Ru[2] * Ru[2];
    c.E12 = Ru[0] * Rv[0] + Ru[1] * Rv[1] + Ru[2] * Rv[2];
    c.E22 = Rv[0] * Rv[0] + Rv[1] * Rv[1] + Rv[2] * Rv[2];
    c.idet = 1.0 / (c.E11 * c.E22 - c.E12 * c.E12);
    double Nv[3];
    cross_d(Ru, Rv, Nv);
    c.iln = 1.0 / sqrt(Nv[0] * Nv[0] + Nv[1] * Nv[1] + Nv[2] * Nv[2]);
#pragma unroll
    for (int k = 0; k < 3; k++) c.n[k] = Nv[k] * c.iln;
    c.L = c.n[0] * Suu[0] + c.n[1] * Suu[1] + c.n[2] * Suu[2];
    c.M = c.n[0] * Suv[0] + c.n[1] * Suv[1] + c.n[2] * Suv[2];
    c.N = c.n[0] * Svv[0] + c.n[1] * Svv[1] + c.n[2] * Svv[2];
    c.T = c.L * c.E22 - 2.0 * c.M * c.E12 + c.N * c.E11;
    c.H = -c.T * c.idet;
    return c;
}

__device__ __forceinline__ void curv_back(const CurvPoint &c, const double Ru[3], const double Rv[3],
                                          const double Suu[3], const double Suv[3], const double Svv[3], double dH,
                                          double D[5][3])
{
    const double dL = -dH * c.E22 * c.idet, dM = 2.0 * dH * c.E12 * c.idet, dN = -dH * c.E11 * c.idet;
    const double ddet = dH * c.T * c.idet * c.idet;
    const double dE11 = -dH * c.N * c.idet + ddet * c.E22;
    const double dE22 = -dH * c.L * c.idet + ddet * c.E11;
    const double dE12 = 2.0 * dH * c.M * c.idet - 2.0 * ddet * c.E12;
    double Q[3];
#pragma unroll
    for (int k = 0; k < 3; k++) Q[k] = dL * Suu[k] + dM * Suv[k] + dN * Svv[k];
    const double qn = Q[0] * c.n[0] + Q[1] * c.n[1] + Q[2] * c.n[2];
    double dNv[3];
#pragma unroll
    for (int k = 0; k < 3; k++) dNv[k] = (Q[k] - qn * c.n[k]) * c.iln;
    double t1[3], t2[3];
    cross_d(Rv, dNv, t1);
    cross_d(dNv, Ru, t2);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        D[0][k] = 2.0 * dE11 * Ru[k] + dE12 * Rv[k] + t1[k];
        D[1][k] = 2.0 * dE22 * Rv[k] + dE12 * Ru[k] + t2[k];
        D[2][k] = dL * c.n[k];
        D[3][k] = dM * c.n[k];
        D[4][k] = dN * c.n[k];
    }
}

// d H / d s along a parameter s of the point: dX holds the derivatives of
// Ru and Rv along s (dRu, dRv) and of the second-derivative rows (dSuu, dSuv,
// dSvv), e.g. for s = u: Ruu, Ruv, Ruuu, Ruuv, Ruvv
__device__ __forceinline__ double curv_dir(const CurvPoint &c, const double Ru[3], const double Rv[3],
                                           const double Suu[3], const double Suv[3], const double Svv[3],
                                           const double dRu[3], const double dRv[3], const double dSuu[3],
                                           const double dSuv[3], const double dSvv[3])
{
    // unit normal along s: P (dRu x Rv + Ru x dRv) / |N|, P = I - n n^T
    double a1[3], a2[3], dn[3];
    cross_d(dRu, Rv, a1);
    cross_d(Ru, dRv, a2);
#pragma unroll
    for (int k = 0; k < 3; k++) dn[k] = a1[k] + a2[k];
    const double pn = dn[0] * c.n[0] + dn[1] * c.n[1] + dn[2] * c.n[2];
#pragma unroll
    for (int k = 0; k < 3; k++) dn[k] = (dn[k] - pn * c.n[k]) * c.iln;
    auto dot = [](const double x[3], const double y[3]) { return x[0] * y[0] + x[1] * y[1] + x[2] * y[2]; };
    const double dE11 = 2.0 * dot(Ru, dRu), dE12 = dot(dRu, Rv) + dot(Ru, dRv), dE22 = 2.0 * dot(Rv, dRv);
    const double dL = dot(dn, Suu) + dot(c.n, dSuu);
    const double dM = dot(dn, Suv) + dot(c.n, dSuv);
    const double dN = dot(dn, Svv) + dot(c.n, dSvv);
    const double dT = dL * c.E22 + c.L * dE22 - 2.0 * (dM * c.E12 + c.M * dE12) + dN * c.E11 + c.N * dE11;
    const double ddet = dE11 * c.E22 + c.E11 * dE22 - 2.0 * c.E12 * dE12;
    return -dT * c.idet + c.T * ddet * c.idet * c.idet;
}

// pgrad for one particle (src/grad.C:3700-3935 regular, 4614-5335
// irregular): E = 1/2 kc a (c1 + c2 - c0)^2 at the particle on the
// alpha-scaled surface, kc a and c0 in de_dr[3p], de_dr[3p + 1].  Writes the
// (u, v) gradient and, per control point, the contribution sum_d t_d dE/dX_d
// (the reduction multiplies by alpha).  X: contractions of unscaled positions
// with the weights of the position and its five derivatives, X3 with the
// third derivatives (uuu, uuv, uvv, vvv).
template <bool IRR>
__device__ void particle_out(const PointArgs &a, int p, int j, const double al[3], const double X[6][3],
                             const double X3[4][3], const double *map, int ncp, const double B[6][12],
                             const double sc[6])
{
    double R[10][3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
#pragma unroll
        for (int d = 1; d < 6; d++) R[d - 1][k] = al[k] * X[d][k];
#pragma unroll
        for (int d = 0; d < 4; d++) R[5 + d][k] = al[k] * X3[d][k];
    }
    const double *Ru = R[0], *Rv = R[1], *Suu = R[2], *Suv = R[3], *Svv = R[4];
    const double *Suuu = R[5], *Suuv = R[6], *Suvv = R[7], *Svvv = R[8];
    const CurvPoint c = curv_setup(Ru, Rv, Suu, Suv, Svv);
    const double kca = a.de_dr[3 * p], pc0 = a.de_dr[3 * p + 1];
    const double dH = kca * (c.H - pc0);
    a.gpuv[2 * p] = dH * curv_dir(c, Ru, Rv, Suu, Suv, Svv, Suu, Suv, Suuu, Suuv, Suvv);
    a.gpuv[2 * p + 1] = dH * curv_dir(c, Ru, Rv, Suu, Suv, Svv, Suv, Svv, Suuv, Suvv, Svvv);
    double D[5][3];
    curv_back(c, Ru, Rv, Suu, Suv, Svv, dH, D);
    auto emit = [&](int x, const double t[6]) {
        double *out = a.contrib + ((size_t)x * a.n + j) * 3;
#pragma unroll
        for (int k = 0; k < 3; k++)
            out[k] = t[1] * D[0][k] + t[2] * D[1][k] + t[3] * D[2][k] + t[4] * D[3][k] + t[5] * D[4][k];
    };
    if constexpr (IRR) {
        for (int x = 0; x < ncp; x++) {
            double t[6];
            irr_weights<6>(map, ncp, x, B, sc, t);
            emit(x, t);
        }
    } else {
#pragma unroll
        for (int x = 0; x < 12; x++) {
            const double t[6] = {B[0][x], B[1][x], B[2][x], B[3][x], B[4][x], B[5][x]};
            emit(x, t);
        }
    }
}

// p: the caller's index of the point, j: its face-sorted position (gradient)
template <bool IRR>
__device__ void point_body(const PointArgs &a, int p, int j, const PointRec &pr, const PointClass &C)
{
    const int nf = C.nf, ncp = IRR ? C.ncp : 12, fc = pr.fc;
    const bool shifted = C.has_img[fc] != 0;
    const double al[3] = {a.r[3 * a.nv], a.r[3 * a.nv + 1], a.r[3 * a.nv + 2]};
    auto load_q = [&](int x, double q[3]) {
        const int v = C.cp[x * nf + fc];
        q[0] = a.r[3 * v];
        q[1] = a.r[3 * v + 1];
        q[2] = a.r[3 * v + 2];
        if (shifted) {
            double o[3];
            image_offset(C.img[x * nf + fc], a.pbc, o);
            q[0] += o[0];
            q[1] += o[1];
            q[2] += o[2];
        }
    };
    double B[6][12];
    point_basis(pr.u, pr.v, B);
    const double sc[6] = {1.0, pr.su, pr.sv, pr.su * pr.su, pr.su * pr.sv, pr.sv * pr.sv};
    const double *map = IRR ? C.map + (size_t)pr.map * 12 * ncp : nullptr;
    (void)map;
    (void)sc;
    // the reference forms the regular gradient's surface from alpha-scaled
    // positions, the irregular one and the evaluation's tangents from
    // unscaled ones (src/grad.C:5822-5850, 6185-6220; src/interpLib.C:8140-8196)
    const bool scale_q = a.grad == 1 && !IRR;
    double X[6][3];
#pragma unroll
    for (int d = 0; d < 6; d++) X[d][0] = X[d][1] = X[d][2] = 0.0;
    auto accumulate = [&](int x, const double t[6]) {
        double q[3];
        load_q(x, q);
        if (scale_q) {
            q[0] *= al[0];
            q[1] *= al[1];
            q[2] *= al[2];
        }
#pragma unroll
        for (int d = 0; d < 6; d++) {
            X[d][0] += t[d] * q[0];
            X[d][1] += t[d] * q[1];
            X[d][2] += t[d] * q[2];
        }
    };
    if constexpr (IRR) {
        for (int x = 0; x < ncp; x++) {
            double t[6];
            irr_weights<6>(map, ncp, x, B, sc, t);
            accumulate(x, t);
        }
    } else {
#pragma unroll
        for (int x = 0; x < 12; x++) {
            const double t[6] = {B[0][x], B[1][x], B[2][x], B[3][x], B[4][x], B[5][x]};
            accumulate(x, t);
        }
    }
    if (a.grad == 2) {
        point_geometry_out(a, p, pr, IRR, al, X);
        return;
    }
    if (a.grad == 4) {
        // the third derivatives for the particle's (u, v) gradient
        double B3[4][12];
        point_basis3(pr.u, pr.v, B3);
        const double sc3[4] = {pr.su * pr.su * pr.su, pr.su * pr.su * pr.sv, pr.su * pr.sv * pr.sv,
                               pr.sv * pr.sv * pr.sv};
        double X3[4][3];
#pragma unroll
        for (int d = 0; d < 4; d++) X3[d][0] = X3[d][1] = X3[d][2] = 0.0;
        for (int x = 0; x < ncp; x++) {
            double t3[4];
            if constexpr (IRR) {
#pragma unroll
                for (int d = 0; d < 4; d++) t3[d] = 0.0;
                for (int y = 0; y < 12; y++) {
                    const double mw = map[y * ncp + x];
#pragma unroll
                    for (int d = 0; d < 4; d++) t3[d] += mw * B3[d][y];
                }
#pragma unroll
                for (int d = 0; d < 4; d++) t3[d] *= sc3[d];
            } else {
#pragma unroll
                for (int d = 0; d < 4; d++) {
                    double s = 0.0;
#pragma unroll
                    for (int y = 0; y < 12; y++) s = (y == x) ? B3[d][y] : s;
                    t3[d] = s;
                }
            }
            double q[3];
            load_q(x, q);
#pragma unroll
            for (int d = 0; d < 4; d++)
#pragma unroll
                for (int k = 0; k < 3; k++) X3[d][k] += t3[d] * q[k];
        }
        particle_out<IRR>(a, p, j, al, X, X3, map, ncp, B, sc);
        return;
    }
    if (!a.grad) {
        double n[3];
        cross_d(X[1], X[2], n);
        const double ln = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        // a zero normal stays zero, as the reference's normalize leaves it
        const double inv = ln > 0.0 ? 1.0 / ln : 0.0;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            a.R[3 * p + k] = X[0][k] * al[k];
            a.nrm[3 * p + k] = n[k] * inv;
        }
        return;
    }
    const double *Ru = X[1], *Rv = X[2], *Suu = X[3], *Suv = X[4], *Svv = X[5];
    const double RuRu = Ru[0] * Ru[0] + Ru[1] * Ru[1] + Ru[2] * Ru[2];
    const double RuRv = Ru[0] * Rv[0] + Ru[1] * Rv[1] + Ru[2] * Rv[2];
    const double RvRv = Rv[0] * Rv[0] + Rv[1] * Rv[1] + Rv[2] * Rv[2];
    const double g = sqrt(RuRu * RvRv - RuRv * RuRv);
    double n[3];
    cross_d(Ru, Rv, n);
    const double ln = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]), iln = 1.0 / ln;
    const double nh[3] = {n[0] * iln, n[1] * iln, n[2] * iln};
    // d nrm / d Ru . s = P (s x Rv) / |n|, d nrm / d Rv . s = P (Ru x s) / |n|
    // with P = I - nh nh^T; the transposes give dE/dRu = (Rv x m) / |n| and
    // dE/dRv = (m x Ru) / |n| for m = P de_dnrm
    auto project = [&](double v[3]) {
        const double d = nh[0] * v[0] + nh[1] * v[1] + nh[2] * v[2];
        v[0] = (v[0] - nh[0] * d) * iln;
        v[1] = (v[1] - nh[1] * d) * iln;
        v[2] = (v[2] - nh[2] * d) * iln;
    };
    const double de_dr[3] = {a.de_dr[3 * p], a.de_dr[3 * p + 1], a.de_dr[3 * p + 2]};
    const double de_dn[3] = {a.de_dnrm[3 * p], a.de_dnrm[3 * p + 1], a.de_dnrm[3 * p + 2]};
    const double frac = a.frac ? a.frac[p] : 1.0;
    double mvec[3] = {de_dn[0], de_dn[1], de_dn[2]};
    project(mvec);
    double de_ru[3], de_rv[3];
    cross_d(Rv, mvec, de_ru);
    cross_d(mvec, Ru, de_rv);
    double Nu[3], Nv[3], t1[3], t2[3];
    cross_d(Suu, Rv, t1);
    cross_d(Ru, Suv, t2);
#pragma unroll
    for (int k = 0; k < 3; k++) Nu[k] = t1[k] + t2[k];
    project(Nu);
    cross_d(Suv, Rv, t1);
    cross_d(Ru, Svv, t2);
#pragma unroll
    for (int k = 0; k < 3; k++) Nv[k] = t1[k] + t2[k];
    project(Nv);
    a.gpuv[2 * p] = (de_dr[0] * Ru[0] + de_dr[1] * Ru[1] + de_dr[2] * Ru[2]) +
                    (de_dn[0] * Nu[0] + de_dn[1] * Nu[1] + de_dn[2] * Nu[2]);
    a.gpuv[2 * p + 1] = (de_dr[0] * Rv[0] + de_dr[1] * Rv[1] + de_dr[2] * Rv[2]) +
                        (de_dn[0] * Nv[0] + de_dn[1] * Nv[1] + de_dn[2] * Nv[2]);
    // per-slot coefficients of the weights t0, t1, t2:
    //   dE/dr_x = frac de_dr t0 + dE/dRu t1 + dE/dRv t2 - on_surface kT / g (dg/dRu t1 + dg/dRv t2)
    const double os = a.on_surface ? a.kT / g : 0.0;
    const double ig = 1.0 / g;
    double c0[3], c1[3], c2[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const double dg_ru = (RvRv * Ru[k] - RuRv * Rv[k]) * ig;
        const double dg_rv = (RuRu * Rv[k] - RuRv * Ru[k]) * ig;
        c0[k] = frac * de_dr[k];
        c1[k] = frac * de_ru[k] - os * dg_ru;
        c2[k] = frac * de_rv[k] - os * dg_rv;
    }
    // slot-major: lanes of a wave write consecutive 24-byte records
    auto emit = [&](int x, const double t[3]) {
        double *out = a.contrib + ((size_t)x * a.n + j) * 3;
#pragma unroll
        for (int k = 0; k < 3; k++) out[k] = c0[k] * t[0] + c1[k] * t[1] + c2[k] * t[2];
    };
    if constexpr (IRR) {
        for (int x = 0; x < ncp; x++) {
            double t[3];
            irr_weights<3>(map, ncp, x, B, sc, t);
            emit(x, t);
        }
    } else {
#pragma unroll
        for (int x = 0; x < 12; x++) {
            const double t[3] = {B[0][x], B[1][x], B[2][x]};
            emit(x, t);
        }
    }
}

// surface::get_pt_coeffs / get_pt_dcoeffs (src/uv_map.C:14-420): the
// control points' vertices and their weights in the position and its
// derivatives along the face's u and v; at the extraordinary vertex the
// reference's mask on the vertex and its ring (derivative weights zero)
template <int KIND>
__device__ void point_weights_out(const PointArgs &a, int p, const PointRec &pr, const PointClass &C)
{
    double *w = a.wts + (size_t)p * 3 * kPointSlots;
    int *id = a.ids + (size_t)p * kPointSlots;
    const int nf = C.nf, fc = pr.fc;
#pragma unroll
    for (int x = 0; x < kPointSlots; x++) {
        w[x] = w[kPointSlots + x] = w[2 * kPointSlots + x] = 0.0;
        id[x] = -1;
    }
    if constexpr (KIND == kPointVertex) {
        const int val = C.valence;
        const double wr = 1.0 / (val * 2);
        id[0] = C.cp[fc];
        w[0] = 0.5;
        for (int x = 1; x <= val; x++) {
            id[x] = C.cp[x * nf + fc];
            w[x] = wr;
        }
    } else {
        double B[6][12];
        point_basis(pr.u, pr.v, B);
        const double sc[6] = {1.0, pr.su, pr.sv, pr.su * pr.su, pr.su * pr.sv, pr.sv * pr.sv};
        const int ncp = KIND == kPointIrregular ? C.ncp : 12;
        const double *map = KIND == kPointIrregular ? C.map + (size_t)pr.map * 12 * ncp : nullptr;
        for (int x = 0; x < ncp; x++) {
            double t[3];
            if constexpr (KIND == kPointIrregular) {
                irr_weights<3>(map, ncp, x, B, sc, t);
            } else {
                // x runs over the 12 slots: pick the basis column without an
                // indexed register array
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    double s = 0.0;
#pragma unroll
                    for (int y = 0; y < 12; y++) s = (y == x) ? B[d][y] : s;
                    t[d] = s;
                }
            }
            id[x] = C.cp[x * nf + fc];
            w[x] = t[0];
            w[kPointSlots + x] = t[1];
            w[2 * kPointSlots + x] = t[2];
        }
    }
}

// one point of kind KIND (j: its position in the batch)
template <int KIND>
__device__ __forceinline__ void point_one(const PointArgs &a, int j)
{
    const int p = (a.grad == 1 || a.grad == 4) ? a.order[j] : j;
    const PointRec pr = a.rec[p];
    const PointClass C = a.cls[pr.cls];
    if (a.grad == 3) {
        point_weights_out<KIND>(a, p, pr, C);
        return;
    }
    if constexpr (KIND == kPointRegular) {
        point_body<false>(a, p, j, pr, C);
    } else if constexpr (KIND == kPointIrregular) {
        point_body<true>(a, p, j, pr, C);
    } else {
        // the extraordinary vertex (evaluation only): 1/2 on the vertex and
        // 1/(2 val) on each ring neighbour at its image next to the vertex,
        // normal unknown (src/interpLib.C:7963-8001)
        const int nf = C.nf, fc = pr.fc, val = C.valence;
        const bool shifted = C.has_img[fc] != 0;
        auto load_q = [&](int x, double q[3]) {
            const int v = C.cp[x * nf + fc];
            q[0] = a.r[3 * v];
            q[1] = a.r[3 * v + 1];
            q[2] = a.r[3 * v + 2];
            if (shifted) {
                double o[3];
                image_offset(C.img[x * nf + fc], a.pbc, o);
                q[0] += o[0];
                q[1] += o[1];
                q[2] += o[2];
            }
        };
        const int i = C.cp[fc];
        const double ri[3] = {a.r[3 * i], a.r[3 * i + 1], a.r[3 * i + 2]};
        double q0[3];
        load_q(0, q0);
        const double w = 1.0 / (val * 2);
        double lr[3] = {0.5 * ri[0], 0.5 * ri[1], 0.5 * ri[2]};
        for (int x = 1; x <= val; x++) {
            double q[3];
            load_q(x, q);
#pragma unroll
            for (int k = 0; k < 3; k++) lr[k] += ((q[k] - q0[k]) + ri[k]) * w;
        }
#pragma unroll
        for (int k = 0; k < 3; k++) {
            a.R[3 * p + k] = lr[k] * a.r[3 * a.nv + k];
            a.nrm[3 * p + k] = __builtin_nan("");
        }
    }
}

// per (face, slot): the face's points summed in order.  Slot-major threads:
// the lanes of a wave take adjacent faces in one slot plane of contrib.
__global__ __launch_bounds__(kPointBlock) void point_segment_kernel(PointReduceArgs a)
{
    const int t = blockIdx.x * kPointBlock + threadIdx.x;
    const int nseg = a.dnseg ? *a.dnseg : a.nseg;
    if (t >= nseg * kPointSlots) return;
    const int x = t / nseg, sidx = t - x * nseg;
    const PointSeg sg = a.seg[sidx];
    const PointClass C = a.cls[sg.cls];
    double *ap = a.apart + (size_t)t * 3;
    if (x >= C.ncp) {
        ap[0] = ap[1] = ap[2] = 0.0;
        return;
    }
    // four records in flight per step, added in order
    double d0 = 0.0, d1 = 0.0, d2 = 0.0;
    const double *c = a.contrib + ((size_t)x * a.n + sg.start) * 3;
    int j = sg.start;
    for (; j + 4 <= sg.end; j += 4, c += 12) {
        double v[12];
#pragma unroll
        for (int i = 0; i < 12; i++) v[i] = c[i];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            d0 += v[3 * i];
            d1 += v[3 * i + 1];
            d2 += v[3 * i + 2];
        }
    }
    for (; j < sg.end; j++, c += 3) {
        d0 += c[0];
        d1 += c[1];
        d2 += c[2];
    }
    const int v = C.cp[x * C.nf + sg.fc];
    double q[3] = {a.r[3 * v], a.r[3 * v + 1], a.r[3 * v + 2]};
    if (C.has_img[sg.fc]) {
        double o[3];
        image_offset(C.img[x * C.nf + sg.fc], a.pbc, o);
        q[0] += o[0];
        q[1] += o[1];
        q[2] += o[2];
    }
    const double *al = a.r + 3 * a.nv;
    double *gb = a.gbuf + (size_t)C.pos[x * C.nf + sg.fc] * 3;
    gb[0] = d0 * al[0];
    gb[1] = d1 * al[1];
    gb[2] = d2 * al[2];
    ap[0] = d0 * q[0];
    ap[1] = d1 * q[1];
    ap[2] = d2 * q[2];
}

// per vertex: its CSR range in order
__global__ __launch_bounds__(kPointBlock) void point_vertex_kernel(PointReduceArgs a)
{
    const int v = blockIdx.x * kPointBlock + threadIdx.x;
    if (v >= a.nv) return;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int e = a.vstart[v]; e < a.vstart[v + 1]; e++) {
        const double *c = a.gbuf + (size_t)e * 3;
        s0 += c[0];
        s1 += c[1];
        s2 += c[2];
    }
    a.gout[3 * v] = s0;
    a.gout[3 * v + 1] = s1;
    a.gout[3 * v + 2] = s2;
}

constexpr int kPointAlphaBlock = 256;

// block b sums its contiguous chunk of src ([len][3]) into dst[b]; the
// partition and the tree depend on len only
__device__ __forceinline__ void point_alpha_block(int len, int chunk, const double *__restrict__ src,
                                                  double *__restrict__ dst)
{
    __shared__ double sh[3][kPointAlphaBlock];
    const int t = threadIdx.x;
    const int lo = blockIdx.x * chunk, hi = min(len, lo + chunk);
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int i = lo + t; i < hi; i += kPointAlphaBlock) {
        s0 += src[3 * (size_t)i];
        s1 += src[3 * (size_t)i + 1];
        s2 += src[3 * (size_t)i + 2];
    }
    sh[0][t] = s0;
    sh[1][t] = s1;
    sh[2][t] = s2;
    __syncthreads();
    for (int h = kPointAlphaBlock / 2; h > 0; h >>= 1) {
        if (t < h) {
            sh[0][t] += sh[0][t + h];
            sh[1][t] += sh[1][t + h];
            sh[2][t] += sh[2][t + h];
        }
        __syncthreads();
    }
    if (t < 3) dst[3 * blockIdx.x + t] = sh[t][0];
}

__global__ __launch_bounds__(kPointAlphaBlock) void point_alpha_kernel(int len, int chunk,
                                                                       const double *__restrict__ src,
                                                                       double *__restrict__ dst)
{
    point_alpha_block(len, chunk, src, dst);
}

// the same two passes with the segment count on the device (pass 0: block b
// of the first partition, blocks past it return; pass 1: one block over the
// first pass's sums)
__global__ __launch_bounds__(kPointAlphaBlock) void point_alpha_dev_kernel(const int *__restrict__ dnseg, int pass,
                                                                           const double *__restrict__ src,
                                                                           double *__restrict__ dst)
{
    const int ns = *dnseg * kPointSlots;
    const int chunk = (ns + kPointAlphaParts - 1) / kPointAlphaParts;
    const int nb = chunk > 0 ? (ns + chunk - 1) / chunk : 0;
    if (pass == 0) {
        if ((int)blockIdx.x >= nb) return;
        point_alpha_block(ns, chunk, src, dst);
    } else {
        point_alpha_block(nb, nb, src, dst);
    }
}

// grad's particle-area term of one face carrying particles (src/grad.C:93-103,
// 212-224, 445-446, 458-459; igrad 6477-6487, 6584, 6819-6820): per Gauss
// point dE/dH = -pfa dAf kc (H - c0_f) / face_area, dAf = g0 w, face_area =
// sum g w of the current (alpha-scaled) surface, no area-element term; the
// face's (face, slot) contributions go to sorted position it.j.
template <int NCP, int NPT, int KIND>
__device__ void density_face(const PointArgs &a, const DensityItem &it, const PointClass &C)
{
    const int nf = C.nf, fc = it.fc;
    const bool shifted = C.has_img[fc] != 0;
    const double al[3] = {a.r[3 * a.nv], a.r[3 * a.nv + 1], a.r[3 * a.nv + 2]};
    double P[NCP][3];
#pragma unroll
    for (int x = 0; x < NCP; x++) {
        const int v = C.cp[x * nf + fc];
        double o[3] = {0.0, 0.0, 0.0};
        if (shifted) image_offset(C.img[x * nf + fc], a.pbc, o);
#pragma unroll
        for (int k = 0; k < 3; k++) P[x][k] = a.r[3 * v + k] + o[k];
    }
    const double *tab = kind_table<KIND, double>();
    const double *wts = tab + NPT * 5 * NCP;
    auto contract = [&](int p, double X[5][3]) {
        const double *T = tab + p * 5 * NCP;
#pragma unroll
        for (int d = 0; d < 5; d++)
#pragma unroll
            for (int k = 0; k < 3; k++) {
                double s = 0.0;
#pragma unroll
                for (int x = 0; x < NCP; x++) s += T[x * 5 + d] * P[x][k];
                X[d][k] = s * al[k];
            }
    };
    double area = 0.0;
    for (int p = 0; p < NPT; p++) {
        double X[5][3], n[3];
        contract(p, X);
        cross_d(X[0], X[1], n);
        area += sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]) * wts[p];
    }
    double G[NCP][3];
#pragma unroll
    for (int x = 0; x < NCP; x++) G[x][0] = G[x][1] = G[x][2] = 0.0;
    for (int p = 0; p < NPT; p++) {
        double X[5][3];
        contract(p, X);
        const CurvPoint c = curv_setup(X[0], X[1], X[2], X[3], X[4]);
        const double dAf = C.g0[p * nf + fc] * wts[p];
        const double dH = -it.pfa * dAf * a.kc * (c.H - it.c0f) / area;
        double D[5][3];
        curv_back(c, X[0], X[1], X[2], X[3], X[4], dH, D);
        const double *T = tab + p * 5 * NCP;
#pragma unroll
        for (int x = 0; x < NCP; x++)
#pragma unroll
            for (int k = 0; k < 3; k++) {
                double s = G[x][k];
#pragma unroll
                for (int d = 0; d < 5; d++) s += T[x * 5 + d] * D[d][k];
                G[x][k] = s;
            }
    }
#pragma unroll
    for (int x = 0; x < NCP; x++) {
        double *out = a.contrib + ((size_t)x * a.n + it.j) * 3;
        out[0] = G[x][0];
        out[1] = G[x][1];
        out[2] = G[x][2];
    }
}

__global__ __launch_bounds__(kPointBlock) void particle_density_kernel(PointArgs a, const DensityItem *items, int count)
{
    const int t = blockIdx.x * kPointBlock + threadIdx.x;
    if (t >= count) return;
    const DensityItem it = items[t];
    const PointClass C = a.cls[it.cls];
    if (C.valence == 6) density_face<12, 15, 0>(a, it, C);
    else if (C.ncp == 11) density_face<11, 16, 1>(a, it, C);
    else density_face<13, 16, 2>(a, it, C);
}

hipError_t launch_particle_density(const PointArgs &a, const DensityItem *items, int count, hipStream_t st)
{
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(particle_density_kernel, dim3((count + kPointBlock - 1) / kPointBlock), dim3(kPointBlock), 0,
                       st, a, items, count);
    return hipGetLastError();
}

// one thread per point of one kind (idx: the batch's positions of that kind)
template <int KIND>
__global__ __launch_bounds__(kPointBlock) void point_kernel(PointArgs a, const int *__restrict__ idx, int count)
{
    const int t = blockIdx.x * kPointBlock + threadIdx.x;
    if (t >= count) return;
    point_one<KIND>(a, idx[t]);
}

// the same with the kinds' counts on the device: kind KIND's indices start
// after the earlier kinds' in idx
template <int KIND>
__global__ __launch_bounds__(kPointBlock) void point_kernel_dev(PointArgs a, const int *__restrict__ idx,
                                                                const int *__restrict__ counts)
{
    const int t = blockIdx.x * kPointBlock + threadIdx.x;
    if (t >= counts[KIND]) return;
    const int start = KIND == 0 ? 0 : (KIND == 1 ? counts[0] : counts[0] + counts[1]);
    point_one<KIND>(a, idx[start + t]);
}

hipError_t launch_points_dev(const PointArgs &a, const int *idx, const int *dcounts, int n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    const dim3 grid((n + kPointBlock - 1) / kPointBlock);
    hipLaunchKernelGGL(point_kernel_dev<kPointRegular>, grid, dim3(kPointBlock), 0, st, a, idx, dcounts);
    hipLaunchKernelGGL(point_kernel_dev<kPointIrregular>, grid, dim3(kPointBlock), 0, st, a, idx, dcounts);
    hipLaunchKernelGGL(point_kernel_dev<kPointVertex>, grid, dim3(kPointBlock), 0, st, a, idx, dcounts);
    return hipGetLastError();
}

// irr_kernel::domain / get_map / get_map_transform (src/irr_kernel.C:
// 380-541), as the host path restates them; the level uses the device's
// log10, which can differ from libm's in the last bit, so a point whose
// -log2(u + v) lies within rounding of an integer may take the adjacent
// level (the sub-patches agree on their shared border)
__device__ int resolve_irregular_dev(double fu, double fv, bool vertex_ok, PointRec &rec)
{
    double f_domain = -log10(fu + fv) / log10(2.0);
    if (!(f_domain < 1e300) || !(f_domain > -1e300)) {
        if (!vertex_ok) return 3;
        if (!(fu < 1e-10 && fv < 1e-10)) return 4;
        rec.kind = kPointVertex;
        rec.map = 0;
        rec.u = rec.v = 0.0;
        rec.su = rec.sv = 1.0;
        return 0;
    }
    if (fu + fv >= 1.0 - 1e-9) f_domain = 1;
    const double dom = ceil(f_domain);
    if (!(dom - 1.0 < (double)kPointDomains)) return 5;
    const int domain = (int)dom;
    const double pow2 = ldexp(1.0, domain - 1);
    double su = pow2, sv = pow2;
    fu *= pow2;
    fv *= pow2;
    int sub;
    if (fu > 0.5) {
        sub = 0;
        fv = 2 * fv;
        fu = 2.0 * fu - 1.0;
        su *= 2.0;
        sv *= 2.0;
    } else if (fv > 0.5) {
        sub = 1;
        fu = 2 * fu;
        fv = 2.0 * fv - 1.0;
        su *= 2;
        sv *= 2;
    } else {
        sub = 2;
        fv = 1 - 2 * fv;
        fu = 1 - 2 * fu;
        su *= -2;
        sv *= -2;
    }
    constexpr double kThresh = 1e-4;   // src/interpLib.C:24
    if (fu + fv > 1.0 + kThresh || fu < -kThresh || fv < -kThresh) return 4;
    rec.kind = kPointIrregular;
    rec.map = (domain - 1) * 3 + sub;
    rec.u = fu;
    rec.v = fv;
    rec.su = su;
    rec.sv = sv;
    return 0;
}

__global__ __launch_bounds__(kPointBlock) void point_resolve_kernel(PointResolveArgs a)
{
    #pragma clang fp contract(off)
    const int i = blockIdx.x * kPointBlock + threadIdx.x;
    if (i >= a.n) return;
    int why = 0;
    PointRec q{};
    const int f = a.faces[i];
    const double u = a.uv[2 * (size_t)i], v = a.uv[2 * (size_t)i + 1];
    if (f < 0 || f >= a.nfaces) {
        why = 1;
    } else {
        const int2 cf = a.face[f];
        q.cls = cf.x;
        q.fc = cf.y;
        q.fu = u;
        q.fv = v;
        if (!isfinite(u) || !isfinite(v)) {
            why = 2;
        } else if (a.cls[cf.x].valence == 6) {
            q.kind = kPointRegular;
            q.map = 0;
            q.u = u;
            q.v = v;
            q.su = q.sv = 1.0;
        } else {
            why = resolve_irregular_dev(u, v, a.vertex_ok != 0, q);
        }
    }
    a.iota[i] = i;
    if (why) {
        a.kind[i] = 3;
        if (a.fkey) a.fkey[i] = a.nfaces;
        atomicMin(a.status, ((unsigned long long)i << 8) | (unsigned long long)why);
        if (a.R)
            for (int k = 0; k < 3; k++) {
                a.R[3 * (size_t)i + k] = __builtin_nan("");
                a.nrm[3 * (size_t)i + k] = __builtin_nan("");
            }
        if (a.geom)
            for (int k = 0; k < kPointGeom; k++) a.geom[(size_t)kPointGeom * i + k] = __builtin_nan("");
        return;
    }
    a.rec[i] = q;
    a.kind[i] = (unsigned char)q.kind;
    if (a.fkey) a.fkey[i] = f;
}

__global__ __launch_bounds__(kPointBlock) void point_seg_flags_kernel(const int *__restrict__ skey, int n, int nfaces,
                                                                      int *__restrict__ flags)
{
    const int j = blockIdx.x * kPointBlock + threadIdx.x;
    if (j >= n) return;
    const int k = skey[j];
    flags[j] = (k < nfaces && (j == 0 || skey[j - 1] != k)) ? 1 : 0;
}

__global__ __launch_bounds__(kPointBlock) void point_seg_fill_kernel(const int *__restrict__ skey,
                                                                     const int *__restrict__ incl, int n, int nfaces,
                                                                     const int2 *__restrict__ face,
                                                                     const PointClass *__restrict__ cls,
                                                                     PointSeg *__restrict__ seg,
                                                                     unsigned char *__restrict__ pkind)
{
    const int j = blockIdx.x * kPointBlock + threadIdx.x;
    if (j >= n) return;
    const int k = skey[j];
    if (k >= nfaces) {
        pkind[j] = 3;
        return;
    }
    const int2 cf = face[k];
    const int sid = incl[j] - 1;
    pkind[j] = cls[cf.x].valence == 6 ? (unsigned char)kPointRegular : (unsigned char)kPointIrregular;
    if (j == 0 || skey[j - 1] != k) {
        seg[sid].cls = cf.x;
        seg[sid].fc = cf.y;
        seg[sid].start = j;
    }
    if (j == n - 1 || skey[j + 1] != k) seg[sid].end = j + 1;
}

hipError_t launch_point_seg_flags(const int *skey, int n, int nfaces, int *flags, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(point_seg_flags_kernel, dim3((n + kPointBlock - 1) / kPointBlock), dim3(kPointBlock), 0, st,
                       skey, n, nfaces, flags);
    return hipGetLastError();
}

hipError_t launch_point_seg_fill(const int *skey, const int *incl, int n, int nfaces, const int2 *face,
                                 const PointClass *cls, PointSeg *seg, unsigned char *pkind, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(point_seg_fill_kernel, dim3((n + kPointBlock - 1) / kPointBlock), dim3(kPointBlock), 0, st,
                       skey, incl, n, nfaces, face, cls, seg, pkind);
    return hipGetLastError();
}

__global__ __launch_bounds__(kPointBlock) void point_accumulate_kernel(const double *__restrict__ src,
                                                                       double *__restrict__ dst, int n, int stride,
                                                                       const unsigned char *__restrict__ skip)
{
    const int i = blockIdx.x * kPointBlock + threadIdx.x;
    if (i >= n) return;
    if (skip && skip[i] == 3) return;
    for (int k = 0; k < stride; k++) dst[(size_t)i * stride + k] += src[(size_t)i * stride + k];
}

hipError_t launch_point_accumulate(const double *src, double *dst, int n, int stride, const unsigned char *skip,
                                   hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(point_accumulate_kernel, dim3((n + kPointBlock - 1) / kPointBlock), dim3(kPointBlock), 0, st,
                       src, dst, n, stride, skip);
    return hipGetLastError();
}

// the points per kind from the kind-sorted keys: the first position of each
// kind, by binary search (one thread)
__global__ void point_counts_kernel(const unsigned char *__restrict__ sorted, int n, int *counts)
{
    if (threadIdx.x != 0) return;
    int first[4];
    for (int k = 0; k < 4; k++) {
        int lo = 0, hi = n;   // first index with key >= k
        while (lo < hi) {
            const int mid = lo + (hi - lo) / 2;
            if (sorted[mid] < k)
                lo = mid + 1;
            else
                hi = mid;
        }
        first[k] = lo;
    }
    for (int k = 0; k < 3; k++) counts[k] = first[k + 1] - first[k];
}

hipError_t launch_point_counts(const unsigned char *sorted, int n, int *counts, hipStream_t st)
{
    hipLaunchKernelGGL(point_counts_kernel, dim3(1), dim3(64), 0, st, sorted, n, counts);
    return hipGetLastError();
}

__global__ void point_resolve_init_kernel(int *counts, unsigned long long *status)
{
    if (threadIdx.x < 3) counts[threadIdx.x] = 0;
    if (threadIdx.x == 0) status[0] = ~0ull;
}

hipError_t launch_point_resolve(const PointResolveArgs &a, hipStream_t st)
{
    hipLaunchKernelGGL(point_resolve_init_kernel, dim3(1), dim3(64), 0, st, a.counts, a.status);
    if (a.n > 0)
        hipLaunchKernelGGL(point_resolve_kernel, dim3((a.n + kPointBlock - 1) / kPointBlock), dim3(kPointBlock), 0,
                           st, a);
    return hipGetLastError();
}

hipError_t launch_points(const PointArgs &a, const int *idx, const int counts[3], hipStream_t st)
{
    const int *ix[3] = {idx, idx + counts[0], idx + counts[0] + counts[1]};
    for (int k = 0; k < 3; k++) {
        if (counts[k] <= 0) continue;
        const dim3 grid((counts[k] + kPointBlock - 1) / kPointBlock);
        if (k == kPointRegular)
            hipLaunchKernelGGL(point_kernel<kPointRegular>, grid, dim3(kPointBlock), 0, st, a, ix[k], counts[k]);
        else if (k == kPointIrregular)
            hipLaunchKernelGGL(point_kernel<kPointIrregular>, grid, dim3(kPointBlock), 0, st, a, ix[k], counts[k]);
        else
            hipLaunchKernelGGL(point_kernel<kPointVertex>, grid, dim3(kPointBlock), 0, st, a, ix[k], counts[k]);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_point_reduce_dev(const PointReduceArgs &a, int nseg_max, hipStream_t st)
{
    hipError_t e = hipMemsetAsync(a.gbuf, 0, sizeof(double) * 3 * (size_t)a.nnz, st);
    if (e != hipSuccess) return e;
    const int ns = nseg_max * kPointSlots;
    if (ns > 0) {
        hipLaunchKernelGGL(point_segment_kernel, dim3((ns + kPointBlock - 1) / kPointBlock), dim3(kPointBlock), 0, st,
                           a);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    hipLaunchKernelGGL(point_vertex_kernel, dim3((a.nv + kPointBlock - 1) / kPointBlock), dim3(kPointBlock), 0, st,
                       a);
    hipLaunchKernelGGL(point_alpha_dev_kernel, dim3(kPointAlphaParts), dim3(kPointAlphaBlock), 0, st, a.dnseg, 0,
                       a.apart, a.partial);
    hipLaunchKernelGGL(point_alpha_dev_kernel, dim3(1), dim3(kPointAlphaBlock), 0, st, a.dnseg, 1, a.partial,
                       a.gout + 3 * (size_t)a.nv);
    return hipGetLastError();
}

hipError_t launch_point_reduce(const PointReduceArgs &a, hipStream_t st)
{
    hipError_t e = hipMemsetAsync(a.gbuf, 0, sizeof(double) * 3 * (size_t)a.nnz, st);
    if (e != hipSuccess) return e;
    const int ns = a.nseg * kPointSlots;
    if (ns > 0) {
        hipLaunchKernelGGL(point_segment_kernel, dim3((ns + kPointBlock - 1) / kPointBlock), dim3(kPointBlock), 0, st,
                           a);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    hipLaunchKernelGGL(point_vertex_kernel, dim3((a.nv + kPointBlock - 1) / kPointBlock), dim3(kPointBlock), 0, st,
                       a);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    // alpha: fixed chunks of apart, then one block over the chunk sums
    const int chunk = (ns + kPointAlphaParts - 1) / kPointAlphaParts;
    const int nb = chunk > 0 ? (ns + chunk - 1) / chunk : 0;
    if (nb > 0) {
        hipLaunchKernelGGL(point_alpha_kernel, dim3(nb), dim3(kPointAlphaBlock), 0, st, ns, chunk, a.apart,
                           a.partial);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        hipLaunchKernelGGL(point_alpha_kernel, dim3(1), dim3(kPointAlphaBlock), 0, st, nb, nb, a.partial,
                           a.gout + 3 * (size_t)a.nv);
    } else {
        e = hipMemsetAsync(a.gout + 3 * (size_t)a.nv, 0, sizeof(double) * 3, st);
        if (e != hipSuccess) return e;
    }
    return hipGetLastError();
}

}  // namespace umcp
