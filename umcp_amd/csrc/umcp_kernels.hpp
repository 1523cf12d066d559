// Kernel argument blocks and launchers shared by the HIP kernels and the host
// surface object.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>

namespace umcp {

constexpr int kFaceBlock = 256;
constexpr int kGatherBlock = 256;

struct PhysParams {
    double kc, kg, KA, k_reg, lipid_c0;
    int mode;
};

// One face class (all faces whose owner vertex has the same valence).
struct FaceArgs {
    const double *r;        // 3 nv + 3 (alpha last)
    int nv;
    int nf;
    const int *cp;          // [ncp][nf]
    const uint8_t *img;     // [ncp][nf]
    const uint8_t *has_img; // [nf]
    const double *g0;       // [npt][nf]
    const double *rurv0;    // [npt][nf]
    const double *Lo, *Li;  // [nf] lipid areas of the face's triangle
    const double *c0;       // [nf]
    const double *table;    // [npt][6][ncp]
    const double *weight;   // [npt]
    double dudv;
    double pbc[3][3];
    double *gslot;          // [3][nslots]
    int slot_base;
    int nslots;
    double *partial;        // [npartial][4]
    int partial_base;
    PhysParams ph;
};

struct GatherArgs {
    int nv;
    const int *vstart;
    const int *vslot;
    const double *gslot;
    int nslots;
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

int face_blocks(int nf, int grp);
hipError_t launch_face_energy_grad(const FaceArgs &a, int ncp, int npt, int grp, hipStream_t st);
hipError_t launch_face_metric(const FaceArgs &a, int ncp, int npt, int mode, double *g0_out,
                              double *rurv0_out, hipStream_t st);
hipError_t launch_vertex_gather(const GatherArgs &a, hipStream_t st);

}  // namespace umcp
