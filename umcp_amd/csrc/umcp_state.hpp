// Internal state of a surface handle, shared by the C-ABI implementation
// (umcp_api.cpp) and the batched ensemble (umcp_ensemble.cpp).
#pragma once

#include <hip/hip_runtime_api.h>

#include <string>
#include <vector>

#include "../../include/umcp.h"
#include "surface_host.hpp"
#include "umcp_lipids.hpp"
#include "umcp_kernels.hpp"

namespace umcp {

struct ClassDev {
    int *cp = nullptr, *pos = nullptr;
    uint16_t *loc = nullptr;
    uint8_t *img = nullptr, *has_img = nullptr;
    double *g0 = nullptr, *rurv0 = nullptr, *ig0 = nullptr, *c0 = nullptr;
    double *lw = nullptr, *ls = nullptr, *lv = nullptr;   // FaceArgs::lw, ls, lv
};

}  // namespace umcp

struct umcp_surface;
namespace umcp {
// per-face curvature-energy density numerators and reference areas in global
// face order, from the face-energy kernel (umcp_api.cpp)
int face_density(umcp_surface *s, const double *r, double *num, double *area0);
}  // namespace umcp

struct umcp_surface {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    umcp::Topology T;
    umcp::Plan P;
    umcp_params prm{};
    bool prepared = false, planned = false, g0_set = false;
    std::vector<double> face_c0;   // global face order
    std::vector<double> face_Lo, face_Li;             // each leaflet's lipid area per face
    // lipid composition (none: one type of prm.lipid_c0 / lipid_apl) and the
    // area of each type per face after setg0, [nfaces][ntypes]
    umcp::LipidComposition lipids;
    std::vector<double> face_lip_o, face_lip_i;
    std::vector<umcp::ClassDev> cls;
    std::vector<void *> allocations;
    size_t device_bytes = 0;              // sum of the allocation sizes
    char *arena = nullptr;                // the plan's arrays (plan_bytes)
    size_t arena_size = 0, arena_used = 0;
    // work buffers
    double *d_r = nullptr, *d_g = nullptr, *d_gbuf = nullptr, *d_partial = nullptr;
    float *d_gbuf32 = nullptr;
    double *d_scalar = nullptr;    // [8]: E, sums[4], ...
    int *d_vstart = nullptr;
    // the face kernel's group records (one per distinct vertex of a 64-face
    // group): slot -> first record, record -> incidence range and list, and
    // the gather's vertex -> records CSR
    int nrecords = 0;
    int *d_gl_start = nullptr;
    int *d_linc_start = nullptr;
    uint16_t *d_linc = nullptr;
    int *d_lvstart = nullptr;
    int *d_lvrec = nullptr;
    int *d_lvert = nullptr;               // [records] vertex of each record
    int npartial = 0;
    // launch shape of the face kernel: waves sharing 64 faces' Gauss points,
    // and whether the faces' control points are staged in LDS
    int point_split = 4;
    int launch_flags = umcp::kLaunchOcc2;
    std::vector<int> launch_order;  // classes in launch order (irregular first)
    std::vector<int> block_start;
    double *d_xi = nullptr;        // host-array BD staging, chunk of steps
    int xi_chunk = 0;
    double *d_energies = nullptr;
    int energies_cap = 0;
    // optional per-launch timing of the face kernels (HIP events on the
    // stream each kernel is launched on)
    bool prof_on = false;
    std::vector<hipEvent_t> prof_events;  // pairs: start, end
    std::vector<int> prof_class;          // class of each pair
    size_t prof_used = 0;
    int prof_every = 1;                   // time one evaluation in prof_every
    long prof_tick = 0;
    // global area / volume restraints
    umcp_restraints rsp{};
    bool has_images = false;              // periodic mesh: no enclosed volume
    double A0_setg0 = 0.0, V_setg0 = 0.0;
    double *d_rs = nullptr;               // [8] dE/dA, dE/dV, E, A, V
    double *d_rpart = nullptr;            // [npartial][2] per-group area and volume
    int *d_fin = nullptr;                 // [1] = 0: the restraint finalize's item (the slot-0 item)
    int *d_nan = nullptr;                 // [1] set by any evaluation whose energy is not finite
    umcp::FaceArgs *d_fa_aux = nullptr;   // a second face table (volume-only passes)
    // Brownian step counter of the device-noise path
    int *d_step = nullptr;
    // the face kernel's work tables: per-class arguments and one item per
    // 64-face group in launch order (re-uploaded when parameters change)
    umcp::FaceArgs *d_fa_tab = nullptr;
    umcp::FaceItem *d_items = nullptr;
    std::vector<umcp::FaceArgs> h_fa;
    bool fa_dirty = true;
    std::vector<umcp::FaceArgs> h_fa_noalloc;   // umcp_energy_grad_dev_noalloc's face table (host copy)
    // momentum-space integrators: truncated inverse effective mass (dense
    // on the device), the step's momenta and the kept product Mi pp
    bool effm_ready = false;
    double effm_point_mass = 0.0;
    double *d_effm = nullptr;             // nv x nv inverse effective mass (row-major), or
    int *d_effm_rowptr = nullptr;                          // its kept entries as CSR
    uint16_t *d_effm_col = nullptr;                        // (16-bit columns: nv <= 65536)
    double *d_effm_val = nullptr;
    double *d_pp_next = nullptr, *d_mp = nullptr;
    long effm_nnz = 0;
    // the scalable form (umcp_effm_build_sparse): M itself in ELL layout and
    // D^-1; products Mi v are Chebyshev solves (residual buffer, two
    // direction buffers, the iteration's coefficients)
    bool effm_solve = false;
    int effm_ell_width = 0;
    int *d_ell_col = nullptr, *d_ell_perm = nullptr;
    double *d_ell_val = nullptr, *d_dinv = nullptr;
    double *d_cheb_x = nullptr, *d_cheb_r = nullptr, *d_cheb_d0 = nullptr, *d_cheb_d1 = nullptr;
    std::vector<double> cheb_c1, cheb_c2;
    double cheb_inv_theta = 0.0, cheb_lo = 0.0, cheb_hi = 0.0;
    std::vector<void *> effm_allocs;
    // per-face energies (umcp_face_energies, umcp_energy_vertex): faces in
    // reference order as (class, index), each vertex's faces as
    // computeModifiedVertices lists them, and the device lists and outputs
    bool fe_ready = false;
    std::vector<int2> fe_items;
    std::vector<int> fe_vstart, fe_vfaces;
    int *d_fe_kinds = nullptr;
    int2 *d_fe_items = nullptr, *d_fe_sub = nullptr;
    double *d_fe_out = nullptr;
    // points on the limit surface (umcp_point_eval / umcp_point_grad): face ->
    // (class, index in the class), the device tables (class table, sub-patch
    // maps), the gradient's (face, slot) buffer and grow-only staging and
    // result buffers
    bool pt_ready = false;
    std::vector<int> pt_face_cls, pt_face_fc;
    umcp::PointClass *d_pt_cls = nullptr;
    int2 *d_pt_face = nullptr;            // per global face: (class, index in the class)
    // device-resident evaluation's work space (grown on demand) and the
    // first rejected point of the last such call
    char *d_ptd = nullptr;
    size_t ptd_cap = 0;
    unsigned long long *d_ptd_status = nullptr;
    double *d_pt_gbuf = nullptr;          // [nnz][3] (face, slot) sums, then the alpha partials
    char *d_pt_in = nullptr, *d_pt_out = nullptr;
    size_t pt_in_cap = 0, pt_out_cap = 0;
    char *pt_stage = nullptr, *pt_host = nullptr;   // pinned (hipHostMalloc)
    size_t pt_stage_cap = 0, pt_host_cap = 0;
    // moving points across faces (umcp_next_face): each face's corner
    // vertices and the face across each of its three edges
    bool nf_ready = false;
    std::vector<int> nf_corner, nf_across;
};

namespace umcp {

constexpr size_t kArenaAlign = 256;
constexpr int kFeSub = 1024;   // faces one vertex's energy may span

// helpers of the surface object used by both translation units
bool restraints_on(const umcp_surface *s);
// device bytes umcp_plan allocates for a prepared surface (the serving
// loop's admission footprint); equals device_bytes after the plan
size_t plan_bytes(const umcp_surface *s);
FaceArgs face_args(umcp_surface *s, size_t c, const double *d_r);

}  // namespace umcp
