// Kernel argument blocks and launchers shared by the HIP kernels and the host
// surface object.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>

namespace umcp {

constexpr int kWave = 64;
constexpr int kGatherBlock = 256;
constexpr int kMetricBlock = 256;
constexpr int kMaxClasses = 4;
constexpr int kMaxSplit = 4;

struct PhysParams {
    double kc, kg, KA, k_reg, lipid_c0;
    int mode;
};

// One face class (all faces whose owner vertex has the same valence).
struct FaceArgs {
    const double *r;        // 3 nv + 3 (alpha last)
    int nv;
    int nf;
    const int *cp;          // [ncp][nf]   control-point vertex ids
    const int *pos;         // [ncp][nf]   position of (face, slot) in the vertex CSR
    const uint8_t *img;     // [ncp][nf]   periodic image codes
    const uint8_t *has_img; // [nf]
    const double *g0;       // [npt][nf]
    const double *rurv0;    // [npt][nf]
    const double *ig0;      // [npt][nf]   1 / g0
    const double *Lo, *Li;  // [nf] lipid areas of the face's triangle
    const double *c0;       // [nf]
    double dudv;
    double pbc[3][3];
    double *gbuf;           // [nnz][3] per-(face, slot) gradient, CSR order (FP64 path)
    float *gbuf32;          // [nnz][3] the same for the mixed path
    PhysParams ph;
};

// All face classes of a surface in one launch: blocks [block_start[c],
// block_start[c+1]) work on class c (kind: 0 = 12 control points / 15 Gauss
// points, 1 = 11 / 16, 2 = 13 / 16).
struct MultiFaceArgs {
    FaceArgs cls[kMaxClasses];
    int kind[kMaxClasses];
    int block_start[kMaxClasses + 1];
    int ncls;
    double *partial;        // [nblocks][4]: energy, dE/dalpha
};

struct GatherArgs {
    int nv;
    const int *vstart;      // nv + 1
    const double *gbuf;     // FP64 contributions, or
    const float *gbuf32;    // FP32 contributions (mixed path) when non-null
    const double *partial;
    int npartial;
    int out_mode;           // 0 gradient, 1 Brownian update, 2 reductions only
    double *g;              // out_mode 0
    double *r;              // out_mode 1
    const double *xi;       // out_mode 1, 3 nv standard normals for this step
    double mobility_dt, noise_amp;
    double *E_out;          // may be null
    double *energies;       // may be null: energies[step] = E
    int step;
    double *sums;           // may be null: the four reduced partial sums
};

int kind_of(int ncp, int npt);
// copies one kind's shared Gauss-point table ([npt][6][ncp]) and weights into
// the device's constant memory (current device)
hipError_t upload_tables(int kind, const double *table, const double *weights, int npt, int ncp);
int face_groups(int nf);
// launch shape: point_split (1..4) waves share the Gauss points of 64 faces
// (whose control points are staged in LDS); flag kLaunchOcc2 bounds registers
// for two waves per SIMD.  feat: energy-form specialisation (features_of).
constexpr int kLaunchOcc2 = 2;
int features_of(const PhysParams &ph);
hipError_t launch_faces(const MultiFaceArgs &a, int nblocks, int point_split, int flags, bool mixed, int feat,
                        hipStream_t st);
hipError_t launch_face_metric(const FaceArgs &a, int kind, int mode, double *g0_out,
                              double *rurv0_out, double *ig0_out, double *partial, int partial_base,
                              hipStream_t st);
int metric_blocks(int nf);
hipError_t launch_vertex_gather(const GatherArgs &a, hipStream_t st);

}  // namespace umcp
