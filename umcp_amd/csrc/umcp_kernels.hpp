// Kernel argument blocks and launchers shared by the HIP kernels and the host
// surface object.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>

namespace umcp {

constexpr int kWave = 64;
constexpr int kGatherBlock = 256;
constexpr int kGatherLanes = 4;                          // lanes summing one vertex's CSR range (4 measured fastest of 2, 4, 8, 16)
constexpr int kGatherVerts = kGatherBlock / kGatherLanes;  // vertices per gather block
constexpr int kGatherUnroll = 8;                          // CSR entries per lane in flight
constexpr int kMetricBlock = 256;
constexpr int kMaxClasses = 4;
constexpr int kMaxSplit = 4;

struct PhysParams {
    double kc, kg, KA, k_reg, lipid_c0;
    int mode;
    int grad_exact;   // Helfrich form: exact dE/dg instead of surface::grad's
};

// One face class (all faces whose owner vertex has the same valence).
struct FaceArgs {
    const double *r;        // 3 nv + 3 (alpha last)
    int nv;
    int nf;
    const int *cp;          // [ncp][nf]   control-point vertex ids
    const uint8_t *img;     // [ncp][nf]   periodic image codes
    const uint8_t *has_img; // [nf]
    const double *g0;       // [npt][nf]
    const double *rurv0;    // [npt][nf]
    const double *ig0;      // [npt][nf]   1 / g0
    // [nf] lipid mode (every type of both leaflets of the face's triangle,
    // j = (type, leaflet), W_j the type's area there, s_j = c0 outer, -c0
    // inner): lw = sum W_j, ls = sum W_j s_j / lw, lv = sum W_j (s_j - ls)^2
    const double *lw, *ls, *lv;
    const double *c0;       // [nf]
    double dudv;
    double pbc[3][3];
    // gradient records of the surface, one per distinct vertex of a 64-face
    // group: group slot s owns records gl_start[s] .. gl_start[s + 1]; record
    // L sums the (face, slot) gradients linc[linc_start[L] .. linc_start[L +
    // 1]) (each slot * 64 + face within the group)
    double *gbuf;           // [records][3] FP64 path
    float *gbuf32;          // [records][3] mixed path
    const int *gl_start;
    const int *linc_start;
    const uint16_t *linc;
    // record L's vertex, and the index among its group's records of the
    // vertex of (face, slot): [ncp][nf] (the face kernels stage a group's
    // control points through these)
    const int *lvert;
    const uint16_t *loc;
    PhysParams ph;
    int *step_ctr;          // Brownian step counter, advanced once per evaluation, or null
    double *partial;        // the surface's block partials [ngroups][4]
    // global restraints (rs == null: none).  rs[0..4] = dE/dA_tot, dE/dV,
    // restraint energy, A_tot, V, written by the restraint finalize; rpart
    // [ngroups][2] the pre-pass's per-group area and volume
    double *rs;
    double *rpart;
    int ngroups;
    // 1: the face kernel's blocks form the restraint scalars from rpart
    // themselves (no finalize launch; the slot-0 block writes rs)
    int rs_fold;
    double KA_g, KV, A0, V0;
    int do_area, do_volume;
};

// One work item of a batched face launch: 64 faces of one class of one
// surface.  slot = the block index the same group has in the surface's own
// launch, so its partial sums land where a launch of that surface alone puts
// them (identical reduction order alone and in a batch).
struct FaceItem {
    int fa, group, slot, kind;
};

struct BatchFaceArgs {
    const FaceArgs *fa;     // device array, one per (surface, class)
    const FaceItem *items;  // device array
    int nitems;
    const double *r;        // non-null: positions for every item (a surface's own launch)
    int *step_ctr;          // non-null: the step counter the slot-0 item advances
    // the restraint finalize: one block per entry, each the index in items of
    // a surface's slot-0 item (surfaces with restraints only)
    const int *fin;
    int nfin;
};

// Counter-based normal deviates: Philox4x64-10 keyed by (seed, 'UMCP') at
// counter (vertex, step, 0, 0), Box-Muller on the first four outputs.
constexpr int kNoiseArray = 0, kNoisePhilox = 1;
constexpr int kEffmBrownian = 0, kEffmLangevin = 1;

struct GatherArgs {
    int nv;
    const int *vstart;      // nv + 1: each vertex's range of vrec
    const int *vrec;        // the records of each vertex (the face kernel's group records)
    const double *gbuf;     // FP64 contributions, or
    const float *gbuf32;    // FP32 contributions (mixed path) when non-null
    const double *partial;
    int npartial;
    int out_mode;           // 0 gradient, 1 Brownian update, 2 reductions only, 3 effective-mass momenta
    double *g;              // out_mode 0
    double *r;              // out_mode 1
    const double *xi;       // out_mode 1, 3 nv standard normals for this step
    double mobility_dt, noise_amp;
    double *E_out;          // may be null
    double *energies;       // may be null: energies[step] = E
    int step;
    double *sums;           // may be null: the four reduced partial sums
    // restraint scalars of the surface (rs[2] = restraint energy) or null
    const double *rs;
    // set to 1 when the evaluation's energy is not finite (may be null)
    int *nan_flag;
    // noise source of out_mode 1
    int noise;              // kNoiseArray (xi) or kNoisePhilox
    unsigned long long seed;
    const int *step_ctr;    // Philox: the step is *step_ctr - 1 (advanced by the face kernel)
    int energies_cap;       // energies[step] is written while step < energies_cap (Philox)
    // out_mode 3: momenta of an effective-mass step (src/hd.C:1546-1693):
    // pp_out = pp_in [+ friction from mp = Mi pp_in] - g Delta/2 + noise - g Delta/2
    const double *pp_in;
    double *pp_out;
    const double *mp;
    int effm_mode;          // kEffmBrownian or kEffmLangevin
    double effm_akma, effm_dt, effm_kT, effm_gamma;
};

// One step of the momentum-space integrators (src/hd.C): the gather
// (out_mode 3) forms the step's momenta from the gradient, then the position
// kernel applies r += (Mi pp) / alpha^2 Delta and keeps Mi pp, the next
// step's friction product.
constexpr int kEffmBlock = 256;
struct EffmArgs {
    int nv;
    // the inverse effective mass: dense (Mi, nv x nv row-major, truncated
    // entries zero) or, when the truncation keeps few enough entries, CSR
    // (rowptr, col, val; columns ascending)
    const double *Mi;
    const int *rowptr;
    const uint16_t *col;    // column indices (nv <= 65536)
    const double *val;
    const double *pp_next;  // momenta of this step (from the gather)
    double *pp_keep;        // momenta carried to the next step (the caller's array)
    double *mp;             // out: Mi pp_keep (Langevin), the next step's friction product
    double *r;              // 3 nv + 3 positions, alpha last
    int mode;               // kEffmBrownian or kEffmLangevin
    double akma, dt;
};
// r += (Mi pp_next) / alpha^2 Delta, pp_keep and mp (one wave per row)
hipError_t launch_effm_position(const EffmArgs &a, hipStream_t st);
// mp = Mi pp (the first step's friction product)
hipError_t launch_effm_product(const EffmArgs &a, const double *pp, hipStream_t st);
// The scalable form: Mi pp by a Chebyshev solve of M q = pp (D^-1
// preconditioned, a fixed number of iterations), the three Cartesian
// components at once, in a reverse Cuthill-McKee order of the vertices
// (perm[p] is the vertex of row p).  Init: r = b (permuted), x = 0,
// d_0 = D^-1 r / theta.  Step k: x += d_k, r -= M d_k,
// d_{k+1} = c1_k d_k + c2_k D^-1 r (reads d_k of the neighbours, writes d_{k+1}
// to the other buffer; four lanes per row split the row's entries).
// Final: out = x + d in vertex order.
constexpr int kChebBlock = 256, kChebLanes = 4, kChebRows = kChebBlock / kChebLanes;
struct ChebArgs {
    int n, width;
    const int *col;         // [width][n] in row order, rows padded with their own index
    const double *val;      // [width][n], padding zero
    const double *dinv;     // 1 / M_pp
    const int *perm;        // vertex of row p
    double *x, *r;          // [n][3] in row order
};
hipError_t launch_cheb_init(const ChebArgs &a, const double *b, double *d0, double inv_theta, hipStream_t st);
hipError_t launch_cheb_step(const ChebArgs &a, const double *d_in, double *d_out, double c1, double c2,
                            hipStream_t st);
hipError_t launch_cheb_final(const ChebArgs &a, const double *d, double *out, hipStream_t st);
// r += (Mi pp_next) / alpha^2 Delta with Mi pp_next already solved into mp,
// and pp_keep (thread per vertex)
hipError_t launch_effm_position_solved(const EffmArgs &a, hipStream_t st);
// set-up of the dense inverse: M from CSR into a zeroed nv x nv array; after
// the in-place potrf + potri (upper triangle, column-major), the full
// symmetric inverse; then the reference's truncation with per-row counts of
// the entries kept
hipError_t launch_effm_scatter(int n, const int *rowptr, const int *col, const double *val, double *A,
                               hipStream_t st);
hipError_t launch_effm_symmetrize(int n, double *A, hipStream_t st);
hipError_t launch_effm_truncate(int n, double *A, double rel_tol, int *rowcount, hipStream_t st);
// the kept entries of the truncated dense array as CSR (rowptr from the
// per-row counts), columns ascending
hipError_t launch_effm_compact(int n, const double *A, const int *rowptr, uint16_t *col, double *val, hipStream_t st);

// ---------------------------------------------------------------------------
// points on the limit surface: surface::evaluateRNRM (src/interpLib.C:
// 7936-8196) and surface::pointGradient (src/grad.C:5697-6443)
// ---------------------------------------------------------------------------
constexpr int kPointDomains = 100;  // sub-patch levels held (MAX_DOMAIN, src/interp.h:20)
constexpr int kPointSlots = 18;     // control points per face at the largest valence
constexpr int kPointBlock = 64;
enum : int { kPointRegular = 0, kPointIrregular = 1, kPointVertex = 2 };

// One point as the host resolves it: the class of its face and the face's
// index there, the kind (regular face, irregular sub-patch, or the
// extraordinary vertex itself), the sub-patch map ((level - 1) * 3 +
// sub-patch) and the coordinates the basis is evaluated at, with the
// derivative scale of the sub-patch (d local / d face parameter; 1 on
// regular faces).
struct PointRec {
    int cls, fc, kind, map;
    double u, v, su, sv;
    double fu, fv;           // the face parameters as given
};

struct PointClass {
    const int *cp;           // [ncp][nf]
    const int *pos;          // [ncp][nf]   position of (face, slot) in the vertex CSR
    const uint8_t *img;      // [ncp][nf]
    const uint8_t *has_img;  // [nf]
    const double *map;       // irregular: [kPointDomains * 3][12][ncp]
    const double *g0;        // [npt][nf] reference metric of the face's Gauss points
    int nf, ncp, valence, pad;
};

// the points of one face, [start, end) of the face-sorted order
struct PointSeg {
    int cls, fc, start, end;
};

struct PointArgs {
    const double *r;         // 3 nv + 3, alpha last
    int nv, n;
    int grad;                // 0: position and normal, 1: gradient, 2: geometry, 3: weights,
                             // 4: particle curvature gradient (pgrad; de_dr[3p] = kc a, c0)
    int on_surface;
    double kT;
    double kc;               // particle gradient: the surface's bending modulus
    double pbc[3][3];
    const PointRec *rec;     // [n], the caller's order
    const int *order;        // gradient: [n] caller index of face-sorted position j
    const PointClass *cls;
    const double *de_dr, *de_dnrm, *frac;  // [n][3], [n][3], [n] (frac null: 1)
    double *R, *nrm;         // position and unit normal [n][3]
    double *contrib;         // dE/dr per (slot, sorted position): [kPointSlots][n][3], no alpha
    double *gpuv;            // (u, v) gradient per point [n][2]
    double *geom;            // geometry per point [n][kPointGeom]
    double *wts;             // weights per point [n][3][kPointSlots] (value, d/du, d/dv)
    int *ids;                // their vertices [n][kPointSlots] (-1 past the last)
};
// geometry record: Ru[3], Rv[3], g, c1 + c2, c1 c2, c1, c2, vec1[2], vec2[2]
constexpr int kPointGeom = 15;
// idx: the positions (eval: caller indices; gradient: face-sorted positions)
// of the regular, irregular and extraordinary-vertex points in that order,
// counts[kind] of each (the vertex kind: evaluation only)
hipError_t launch_points(const PointArgs &a, const int *idx, const int counts[3], hipStream_t st);
// grad's particle-area term of one face (src/grad.C:445-446, 458-459,
// 6819-6820): its (face, slot) contributions land at sorted position j of the
// face's segment, after the face's particles
struct DensityItem {
    int cls, fc, j, pad;
    double pfa;              // particle area of the face's particles whose c0 differs from the face's
    double c0f;              // the face's spontaneous curvature
};
hipError_t launch_particle_density(const PointArgs &a, const DensityItem *items, int count, hipStream_t st);
// gradient reduction of a face-sorted batch: per (face, slot) the sum of its
// points' contributions in order, alpha-scaled into gbuf at the slot's vertex
// CSR position (gbuf zeroed first), and its alpha-gradient part into apart;
// then per vertex the sum of its CSR range in order into gout [nv][3], and
// the alpha gradient into gout[nv] by a two-stage tree over apart
// Device-resident evaluation (umcp_point_eval_dev): each point resolved on
// the device by the host path's rules (face -> class and index, the
// irregular sub-patch level and local coordinates, the same rejections);
// kind[i] is the point's kind (3: rejected, whose outputs become NaN), iota
// the point indices for the stable partition by kind; status holds the first rejected point as (index << 8 | reason),
// reasons: 1 face out of range, 2 non-finite (u, v), 3 the extraordinary
// vertex without a map, 4 outside the face, 5 level beyond MAX_DOMAIN.
struct PointResolveArgs {
    int n, nfaces, vertex_ok;
    const int *faces;
    const double *uv;
    const int2 *face;        // per global face: (class, index in the class)
    const PointClass *cls;
    PointRec *rec;
    unsigned char *kind;
    int *iota;
    int *counts;             // [3], zeroed here (the partition's counts fill them)
    unsigned long long *status;
    double *R, *nrm;         // evaluation: rejected points get NaN (null otherwise)
    double *geom;            // geometry: rejected points get NaN records (null otherwise)
    int *fkey;               // gradient: the face sort's keys (nfaces for a rejected point), or null
};
// counts zeroed and status set to all ones first (a fill kernel)
hipError_t launch_point_resolve(const PointResolveArgs &a, hipStream_t st);
// counts[k] = the points of kind k, from the kind-sorted keys
hipError_t launch_point_counts(const unsigned char *sorted, int n, int *counts, hipStream_t st);
// the point kernels over the kind-partitioned indices with the per-kind
// counts on the device (grids sized for n points, threads past a kind's
// count return)
hipError_t launch_points_dev(const PointArgs &a, const int *idx, const int *dcounts, int n, hipStream_t st);
// A gradient batch on the device: after the stable sort by face (skey the
// sorted face keys), the segments (one per face with points: incl the
// inclusive scan of the head flags, so a position's segment is incl - 1 and
// the count is incl[n - 1]) and each sorted position's kind (3: rejected)
hipError_t launch_point_seg_flags(const int *skey, int n, int nfaces, int *flags, hipStream_t st);
hipError_t launch_point_seg_fill(const int *skey, const int *incl, int n, int nfaces, const int2 *face,
                                 const PointClass *cls, PointSeg *seg, unsigned char *pkind, hipStream_t st);
// dst[i] += src[i] (i < n), skipping i with skip[i] == 3 when skip is given
// (pairs: stride 2 entries per i)
hipError_t launch_point_accumulate(const double *src, double *dst, int n, int stride, const unsigned char *skip,
                                   hipStream_t st);
// radix sort of int keys (end_bit bits) carrying int values, stable; the
// inclusive sum of int flags
hipError_t point_sort_by_face(const int *keys_in, int *keys_out, const int *vals_in, int *vals_out, int n, int end_bit,
                              void *temp, size_t *temp_bytes, hipStream_t st);
hipError_t point_inclusive_sum(const int *in, int *out, int n, void *temp, size_t *temp_bytes, hipStream_t st);
// stable partition by kind (radix sort of the 2-bit keys with the indices as
// values); temp == nullptr sets *temp_bytes to the work space needed
hipError_t point_partition_by_kind(const unsigned char *keys_in, unsigned char *keys_out, const int *vals_in,
                                   int *vals_out, int n, void *temp, size_t *temp_bytes, hipStream_t st);

struct PointReduceArgs {
    const double *r;
    int nv, n, nseg, nnz;
    double pbc[3][3];
    const PointSeg *seg;     // [nseg]
    const PointClass *cls;
    const double *contrib;
    const int *vstart;       // [nv + 1]
    double *gbuf;            // [nnz][3]
    double *apart;           // [kPointSlots][nseg][3]
    double *partial;         // [kPointAlphaParts][3]
    double *gout;            // [nv + 1][3]
    const int *dnseg;        // device-resident batches: the segment count on the device
};
constexpr int kPointAlphaParts = 256;
hipError_t launch_point_reduce(const PointReduceArgs &a, hipStream_t st);
// the reduction with the segment count on the device (a.dnseg; grids for
// nseg_max segments); the alpha partition and tree as launch_point_reduce's
hipError_t launch_point_reduce_dev(const PointReduceArgs &a, int nseg_max, hipStream_t st);

// energies of listed faces (surface::faceEnergy, src/interpLib.C:3405-3424):
// item i = (class, index in the class) of the face whose energy lands in E[i],
// its area (sum dudv g w) in E[n + i], reference area (sum dudv g0 w) in
// E[2n + i] and the numerator of its curvature-energy density (penergy's
// face_energy_density) in E[3n + i]; fa: the surface's per-class face
// tables, r its positions
hipError_t launch_face_energies(const FaceArgs *fa, const int *kinds, int nclass, const double *r, const int2 *items,
                                int n, double *E, hipStream_t st);

struct GatherItem {
    int surf, vblock;
};

struct BatchGatherArgs {
    const GatherArgs *ga;   // device array, one per surface
    const GatherItem *items;
    int nitems;
};

int kind_of(int ncp, int npt);
// copies one kind's shared Gauss-point table ([npt][6][ncp]), weights and
// corner weights ([3][ncp]) into the device's constant memory (current device)
hipError_t upload_tables(int kind, const double *table, const double *weights, const double *corner, int npt,
                         int ncp);
int face_groups(int nf);
// launch shape: point_split (1..4) waves share the Gauss points of 64 faces
// (whose control points are staged in LDS); flag kLaunchOcc2 bounds registers
// for two waves per SIMD.  feat: energy-form specialisation (features_of).
constexpr int kLaunchOcc2 = 2;
// flag kLaunchMfma (mixed precision only): the basis contractions run on the
// matrix cores (faces_batch_mfma; point splits 1, 2 and 4)
constexpr int kLaunchMfma = 4;
// flag kLaunchCps (point splits 2..4): the waves of a group also split the
// control points -- each wave reverses its own Gauss points into LDS and
// accumulates the gradient of its share of the control points over every
// point (faces_batch_kernel's CPS form); with kLaunchOcc2 the registers are
// bounded for three waves per SIMD
constexpr int kLaunchCps = 8;
constexpr int kLaunchFlags = kLaunchOcc2 | kLaunchMfma | kLaunchCps;
// null if (point_split, flags) names a launch shape, else the reason it does not
inline const char *launch_shape_error(int point_split, int flags)
{
    if (point_split < 1 || point_split > kMaxSplit) return "point split must be between 1 and 4 waves";
    if (flags & ~kLaunchFlags) return "unknown launch flags";
    if ((flags & kLaunchMfma) && (flags & kLaunchCps)) return "the matrix-core path has no control-point split";
    if ((flags & kLaunchMfma) && point_split == 3) return "the matrix-core path takes point split 1, 2 or 4";
    if ((flags & kLaunchCps) && point_split == 1) return "the control-point split takes point split 2, 3 or 4";
    return nullptr;
}
int features_of(const PhysParams &ph);
hipError_t launch_face_metric(const FaceArgs &a, int kind, int mode, double *g0_out,
                              double *rurv0_out, double *ig0_out, double *partial, int partial_base,
                              hipStream_t st);
int metric_blocks(int nf);
hipError_t launch_vertex_gather(const GatherArgs &a, hipStream_t st);
int gather_blocks(int nv);

// restraint pre-pass over the same items as the face kernel (per-group area
// and volume), then the finalize for the surfaces a.fin names: per surface,
// the partials summed in fixed order into the restraint scalars.  Surfaces of
// at most kFoldMaxGroups groups fold the finalize into the face kernel
// instead: each face block loads the partials beside its group's staging
// loads and sums them itself, in the finalize's order.  On the 1281-group
// vesicle the fold costs the face kernel 3.6 us against the finalize
// launch's 2.6 (4.2 us kernel), so large surfaces keep the launch.
constexpr int kFoldMaxGroups = 256;
hipError_t launch_restraint_prepass(const BatchFaceArgs &a, hipStream_t st);

// face kernel over a work-item table (one surface or many); grid = 0: one
// workgroup per item, else a persistent grid of that many workgroups striding
// over the items.  Batched gather and restraint launches likewise.
hipError_t launch_faces_batch(const BatchFaceArgs &a, int grid, int point_split, int flags, bool mixed, int feat,
                              hipStream_t st);
hipError_t launch_gather_batch(const BatchGatherArgs &a, int grid, bool mixed, hipStream_t st);

// 3 n normal deviates of (seed, step) for vertices [0, n) (the stream the
// Brownian update draws from)
hipError_t launch_noise(unsigned long long seed, int step, int n, double *out, hipStream_t st);

}  // namespace umcp
