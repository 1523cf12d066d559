/*
 * umcp.h -- C ABI of libumcp_core.so, the MI355X (gfx950) implementation of
 * the UMCP membrane-surface hot path: the Loop-subdivision limit-surface
 * Helfrich energy, its vertex gradient and the Brownian step that consumes it.
 *
 * Each entry point replaces one `surface` method or driver step of the
 * reference (alexsodt/UMCP); the comment above each gives the interface it
 * replaces.  Host arrays use the reference layout: r and g are double[3nv+3],
 * x y z per vertex followed by the box-scale triple alpha (src/hd.C:232-247).
 * Device state (mesh, tables, work buffers) lives in HBM behind the handle.
 * All calls return UMCP_OK or a negative status; umcp_last_error() gives text.
 * Calls on one handle are not thread safe; distinct handles are independent.
 */
#ifndef UMCP_H
#define UMCP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define UMCP_OK 0
#define UMCP_ERR_BAD_MESH (-1)      /* src/interpLib.C:1287-1291, 5500-5509 exit(1) */
#define UMCP_ERR_NAN_ENERGY (-2)    /* src/interpLib.C:3546-3550 "nan en." exit(1) */
#define UMCP_ERR_OOM (-3)
#define UMCP_ERR_DEVICE (-4)
#define UMCP_ERR_ARG (-5)
#define UMCP_ERR_STATE (-6)         /* e.g. grad before plan, src/grad.C:67-72 */

/* energy forms */
#define UMCP_MODE_LIPID 0     /* reference default: FIXED_A + LOCAL_LIPID_ENERGY (src/local_config.h:21-22) */
#define UMCP_MODE_HELFRICH 1  /* per-face c0, g0 area element (#else branches of src/interpLib.C:3766-3844) */

/* arithmetic of the per-face kernels */
#define UMCP_PREC_FP64 0
#define UMCP_PREC_MIXED 1     /* FP32 per-face arithmetic, FP64 reductions and positions */

/* gradient of the Helfrich form (UMCP_MODE_HELFRICH; the lipid form has one
 * gradient).  UMCP_GRAD_REFERENCE (default) is what surface::grad computes:
 * on regular faces dE/dg also carries 1/2 kc (c1+c2-c0)^2 dudv w
 * (src/grad.C:448) although the energy's area element is the fixed g0;
 * irregular faces carry no such term (igrad keeps it under #ifndef FIXED_A,
 * src/grad.C:6822-6824).  UMCP_GRAD_EXACT is the exact derivative of the
 * energy instead. */
#define UMCP_GRAD_REFERENCE 0
#define UMCP_GRAD_EXACT 1

typedef struct umcp_surface umcp_surface;

/* Physical constants of one surface; replaces the process globals kc, kg,
 * KA, k_reg (src/interpLib.C:85-95, set by src/hd.C:125-128) and the lipid
 * library entry "ZERO" (src/lipid_composition.C:18-34). Energies kcal/mol,
 * lengths Angstrom. */
typedef struct {
    double kc;
    double kg;
    double KA;
    double k_reg;
    int mode;
    int precision;
    double lipid_c0;
    double lipid_apl;
    int grad_form;         /* UMCP_GRAD_REFERENCE (0) or UMCP_GRAD_EXACT */
} umcp_params;

/* Per-vertex overdamped Brownian step, r <- r - mobility_dt * dE/dr + noise_amp * xi
 * with mobility_dt = D dt / kT and noise_amp = sqrt(2 D dt) (docs/timescales.tex:59-67). */
typedef struct {
    double mobility_dt;
    double noise_amp;
} umcp_bd_params;

/* link probe; replaces hd_umcp_link_check (src/interpLib.C:9577-9582) */
int umcp_link_check(void);

/* selects the HIP device for handles created afterwards on this thread */
int umcp_set_device(int device);

/* surface::loadLattice(fileName, noise=0) (src/interp.h:291): parse the .mesh
 * file and orient it as the reference does */
int umcp_load_mesh(const char *path, umcp_surface **out);
void umcp_destroy(umcp_surface *s);
const char *umcp_last_error(const umcp_surface *s);

int umcp_nv(const umcp_surface *s);
/* number of faces (regular + irregular); nreg, nirr may be NULL */
int umcp_nfaces(const umcp_surface *s, int *nreg, int *nirr);

/* surface::get (src/interp.h:427): positions of the loaded mesh, alpha = 1 */
int umcp_get(const umcp_surface *s, double *r);

/* parameters; see umcp_params */
int umcp_set_params(umcp_surface *s, const umcp_params *p);

/* surface::readLipidComposition(FILE *) (src/interp.h:536,
 * src/lipid_composition.C:47-333): the reference's lipid library extended by
 * the file's "lipid library NAME APL c0" lines and the composition of its
 * "lipid inner NAME parts" / "lipid outer NAME parts" lines; the next
 * umcp_setg0 packs each triangle's leaflets with the reference's rand()
 * draws (src/interpLib.C:3127-3196, 3286-3358) from a glibc-compatible
 * stream seeded with seed (the reference's srand(random_seed), hd.C:118).
 * The lipid-mode energy then sums over the types (src/interpLib.C:3818-3838).
 * Without a composition: one type of umcp_params' lipid_c0 and lipid_apl.
 * UMCP_ERR_ARG for an unreadable file, a malformed directive or a lipid not
 * in the library (the reference exits or reads out of bounds). */
int umcp_read_lipid_composition(umcp_surface *s, const char *path, unsigned int seed);
/* the same composition from arrays: ntypes (1..64) types with spontaneous
 * curvature c0 (1/A), area per lipid apl (A^2) and input parts per leaflet
 * (inner_parts, outer_parts >= 0, each summing above 0) */
int umcp_set_lipids(umcp_surface *s, int ntypes, const double *c0, const double *apl,
                    const double *inner_parts, const double *outer_parts, unsigned int seed);
/* number of lipid types in force (1 without a composition) */
int umcp_lipid_type_count(const umcp_surface *s);

/* host half of surface::generatePlan (src/interp.h:337): face ownership,
 * control-point lists, periodic image codes and the shared Gauss-point
 * tables; needs no GPU (introspection works after it) */
int umcp_prepare(umcp_surface *s);

/* surface::generatePlan (src/interp.h:337): umcp_prepare, then the plan,
 * per-face state and work buffers resident on the device */
int umcp_plan(umcp_surface *s);

/* a new handle with the prepared host plan of tmpl (the reference's
 * copyFrom argument of loadLattice, src/interp.h:291): no parsing,
 * orientation or table construction; the copy is then planned, parametrised
 * and setg0'd on its own (on the current device) */
int umcp_clone(const umcp_surface *tmpl, umcp_surface **out);

/* device bytes of a prepared surface: what umcp_plan allocates (planned, also
 * before the plan) and what this handle holds now (allocated) */
int umcp_device_bytes(const umcp_surface *s, size_t *planned, size_t *allocated);

/* per-face spontaneous curvature (UMCP_MODE_HELFRICH), reference face order
 * (regular faces, then irregular); replaces theVertices[i].c0 of the owner */
int umcp_set_face_c0(umcp_surface *s, const double *c0);

/* surface::setg0(r) (src/interp.h:293): reference metric, RuRv0 and the
 * single-lipid packing per triangle */
int umcp_setg0(umcp_surface *s, const double *r);

/* surface::energy(r, NULL) (src/interp.h:294) */
int umcp_energy(umcp_surface *s, const double *r, double *E);

/* surface::grad(r, g) (src/interp.h:304): ADDS dE/dr (and dE/dalpha) into g */
int umcp_grad(umcp_surface *s, const double *r, double *g);

/* energy and gradient in one pass; g is overwritten (not accumulated) */
int umcp_energy_grad(umcp_surface *s, const double *r, double *E, double *g);

/* surface::area (src/interp.h:301): geometric area sum(dudv g w) and
 * reference area sum(dudv g0 w) */
int umcp_area(umcp_surface *s, const double *r, double *A, double *A0);
/* Energies of single faces and of the faces a vertex moves: umcp_face_energies
 * writes surface::faceEnergy (src/interpLib.C:3405-3424) of every face, in the
 * reference's face order, into E[nfaces], and (when not NULL) each face's
 * terms of surface::area into A (sum dudv g w) and A0 (sum dudv g0 w);
 * umcp_energy_vertex replaces
 * surface::energy(r, NULL, do_vertex) with do_vertex >= 0
 * (src/interpLib.C:4262-4295): the energy of every face that has the vertex
 * among its control points, summed in computeModifiedVertices' order (faces
 * ascending, one term per occurrence), without the global restraint terms. */
int umcp_face_energies(umcp_surface *s, const double *r, double *E, double *A, double *A0);
int umcp_energy_vertex(umcp_surface *s, const double *r, int vertex, double *E);
/* surface::area(r, do_vertex, &A, &A0) with do_vertex >= 0
 * (src/interpLib.C:4420-4560): area and reference area of the same faces */
int umcp_area_vertex(umcp_surface *s, const double *r, int vertex, double *A, double *A0);

/* nsteps Brownian steps from host arrays; xi is [nsteps][3nv] standard
 * normals; r (3nv+3) is updated in place; energies (nsteps, may be NULL)
 * receives the energy at the start of each step (src/hd.C:1128, 1543-1720) */
int umcp_bd_steps(umcp_surface *s, double *r, const double *xi, int nsteps,
                  const umcp_bd_params *bd, double *energies);

/* Device-resident variants: all pointers are device pointers (HBM); stream
 * is a hipStream_t (NULL = the surface's own stream).  They enqueue work and
 * return without synchronising, so they cannot return UMCP_ERR_NAN_ENERGY
 * themselves: every evaluation whose energy is not finite (the reference's
 * "nan en." exit, src/interpLib.C:3546-3550) sets a flag on the surface, and
 * umcp_nan_check reports it. */
int umcp_energy_grad_dev(umcp_surface *s, const double *d_r, double *d_E, double *d_g,
                         void *stream);
int umcp_bd_steps_dev(umcp_surface *s, double *d_r, const double *d_xi, int nsteps,
                      const umcp_bd_params *bd, double *d_energies, void *stream);
/* Waits for the device, then returns UMCP_ERR_NAN_ENERGY if any evaluation
 * of s since the last report (by this call or by a host entry point that
 * returned UMCP_ERR_NAN_ENERGY) had a non-finite energy, and clears the
 * report; UMCP_OK otherwise.  No counterpart in the reference, which exits. */
int umcp_nan_check(umcp_surface *s);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under stream capture, serving loops that forbid allocation):
 * umcp_energy_grad_dev on a planned, setg0'd surface, calling no hipMalloc or
 * hipFree.  Its work buffers (face tables, gradient records, block and
 * restraint partials) live in one static device array of
 * UMCP_NOALLOC_SCRATCH_BYTES per device instead of the handle's buffers;
 * results are bitwise those of umcp_energy_grad_dev.  A surface whose buffers
 * do not fit (about three records of 3 doubles per vertex: roughly 110k
 * vertices in FP64, twice that mixed) returns UMCP_ERR_ARG.  All calls on one
 * device share its array and keep the last call's contents: calls that
 * overlap on one device (other streams or threads) race. */
#define UMCP_NOALLOC_SCRATCH_BYTES ((size_t)8 << 20)
int umcp_energy_grad_dev_noalloc(umcp_surface *s, const double *d_r, double *d_E, double *d_g,
                                 void *stream);

/* Global restraints: area KA_global*A0*((A-A0)/A0)^2 with A = sum dudv g w
 * (the GLOBAL_AREA form, src/interpLib.C:4322-4326) and volume
 * KV/2*V0*((V-V0)/V0)^2 (src/minimize.C:90-114) of the limit-point
 * polyhedron (src/interpLib.C:997-1098).  A0 <= 0 takes the reference area
 * sum(dudv g0 w) of setg0, V0 == 0 the volume at setg0 (src/hd.C:273-279).
 * Once set they are part of every energy, gradient and Brownian step.  KV
 * needs a closed mesh (no periodic images).  KA_global in kcal/mol/A^2, KV in
 * kcal/mol/A^3. */
typedef struct {
    double KA_global;
    double KV;
    double A0;
    double V0;
} umcp_restraints;
int umcp_set_restraints(umcp_surface *s, const umcp_restraints *rs);

/* surface::volume (src/interpLib.C:997-1098): volume enclosed by the
 * polyhedron of the limit positions of the vertices (closed meshes) */
int umcp_volume(umcp_surface *s, const double *r, double *V);

/* Momentum-space integrators of src/hd.C with the effective-mass matrix
 * (Cartesian momenta, one per vertex coordinate).
 * umcp_effm_build (getSparseEffectiveMass, src/modes.C:1719-1960, over the
 * 66-point lattice of every face and the point mass of src/hd.C:418-481;
 * after umcp_setg0, which fixes the reference area): the mass matrix (host,
 * sparse), its Cholesky inverse on the device (rocSOLVER potrf + potri, the
 * reference's LAPACK pair), truncated as the reference does (rel_tol
 * 1e-20) and kept dense in HBM (nv x nv doubles; nv <= UMCP_EFFM_MAX_NV).
 * umcp_effm_steps(_dev): per step g = grad(r); UMCP_EFFM_BD
 * (do_bd_membrane): pp = -g Delta + gamma^-1 N(0, s) / Delta, momenta cleared
 * after the step; UMCP_EFFM_LD (do_ld, OLD_LANGEVIN friction): pp += -Delta
 * gamma (Mi pp) / alpha^2 - g Delta + N(0, s), momenta carried; then r +=
 * (Mi pp) / alpha^2 Delta, with Delta = AKMA_TIME dt and s = sqrt(2 gamma kT
 * Delta) (src/hd.C:1543-1720, 1943-1947).  xi: [nsteps][3nv] standard
 * normals in (vertex, coordinate) order, or NULL (device entry only) for the
 * device Philox stream of (seed, step counter) that umcp_bd_steps_rng_dev
 * uses. */
#define UMCP_EFFM_BD 0
#define UMCP_EFFM_LD 1
#define UMCP_EFFM_MAX_NV 65536   /* umcp_effm_build (dense inverse) only */
typedef struct {
    int mode;              /* UMCP_EFFM_BD or UMCP_EFFM_LD */
    double dt;             /* time step, s */
    double kT;             /* kcal/mol */
    double gamma;          /* gamma_langevin, the reference's units */
} umcp_effm_params;
int umcp_effm_build(umcp_surface *s, long *nnz, double *point_mass);
/* the (truncated) inverse effective mass as a dense nv x nv array; *nnz of
 * umcp_effm_build counts the entries the truncation keeps */
int umcp_effm_get(umcp_surface *s, double *Minv);
/* The scalable form of getSparseEffectiveMass (src/modes.C:1719-1960): the
 * mass matrix M stays sparse on the device (ELL, about 37 entries per row)
 * and every product Mi v of the integrators is a solve of M x = v by a
 * Chebyshev iteration preconditioned with diag(M)^-1, of a fixed length set
 * here from the spectral interval of diag(M)^-1 M (Gershgorin above, Lanczos
 * below) and checked on a pseudo-random right-hand side: the relative
 * residual is below tol (0: 1e-12; accepted 1e-14 .. 1e-3).  The reference's
 * truncation drops entries below 1e-20 of the diagonal, so the two forms'
 * products agree to the solve's tolerance.  No vertex limit; memory O(nv).
 * *nnz: the entries of M; *iterations: the solve's length.  Replaces an
 * earlier umcp_effm_build (and the reverse); umcp_effm_get then fails
 * (UMCP_ERR_STATE), umcp_effm_apply gives products. */
int umcp_effm_build_sparse(umcp_surface *s, double tol, long *nnz, double *point_mass, int *iterations);
/* out = Mi v for the built form (host arrays [nv][3]): the product with the
 * truncated inverse (umcp_effm_build) or the solve (umcp_effm_build_sparse) */
int umcp_effm_apply(umcp_surface *s, const double *v, double *out);
/* host arrays: r (3nv+3) and pp (3nv) in/out, xi [nsteps][3nv], energies
 * (energy at the start of each step) may be NULL */
int umcp_effm_steps(umcp_surface *s, double *r, double *pp, const double *xi, int nsteps,
                    const umcp_effm_params *p, double *energies);
/* device arrays on the surface's device; stream may be NULL */
int umcp_effm_steps_dev(umcp_surface *s, double *d_r, double *d_pp, const double *d_xi, int nsteps,
                        const umcp_effm_params *p, unsigned long long seed, double *d_energies, void *stream);

/* On-disk formats.  umcp_save_mesh: surface::saveSurface
 * (src/write_structures.C:473-540) with positions r (3nv+3; NULL: the
 * file's).  Restart files (Simulation::saveRestart / loadRestart,
 * src/restart.C:7-168): the alpha line, nv + 1 rows of positions (the last is
 * the alpha row) with the momenta when pp is given (six values per row, as
 * the reference writes for Cartesian momenta), then "seed N".  Values are
 * written with %.17g; the reader also takes the reference's %lf files, skips a
 * complex section and reports has_pp (six-value rows) and seed (-1: none). */
int umcp_save_mesh(umcp_surface *s, const double *r, const char *path);
int umcp_write_restart(const char *path, int nv, const double *r, const double *pp, int seed);
int umcp_read_restart(const char *path, int nv, double *r, double *pp, double *alpha, int *has_pp, int *seed);
/* trajectory frame (surface::writeLimitingSurface, src/write_structures.C:
 * 662-723): XYZ, "C x y z" for the limit positions at the corners (0,0),
 * (0,1), (1,0) of every face in the reference's face order (the reference's
 * plim = 1), scaled by alpha; append != 0 adds a frame to the file.  Needs
 * umcp_prepare. */
int umcp_write_limit_surface(umcp_surface *s, const double *r, const char *path, int append);
/* the topology of those frames (surface::writeLimitingSurfacePSF,
 * src/write_structures.C:544-655 over writePSF, src/mutil.C:601-676, without
 * complexes): PSF EXT, one "C" atom per frame point, three bonds per face
 * joining its corners.  Needs umcp_prepare. */
int umcp_write_limit_psf(umcp_surface *s, const char *path);

/* Points on the limit surface, batched; faces in the reference's numbering
 * (regular faces first, then irregular), (u, v) the face's parameters.
 * umcp_point_eval replaces surface::evaluateRNRM (src/interpLib.C:7936-8196):
 * R[3i..] = alpha-scaled position, nrm[3i..] = unit normal of the unscaled
 * tangents; at an irregular face's extraordinary vertex (u = v = 0) the
 * reference's mask (1/2 on the vertex, 1/(2 val) on each neighbour) and a NaN
 * normal.  umcp_point_grad replaces a loop of surface::pointGradient
 * (src/grad.C:5697-6443) over the points: for each point i it adds to
 * g (3nv+3) the vertex and alpha gradient of an energy with derivatives
 * de_dr[3i..] and de_dnrm[3i..] at the point (times frac_mult[i]; NULL: 1),
 * plus -kT dlog(g)/dr when on_surface, and adds to g_puv[2i..] its derivative
 * in (u, v).  Like the reference, the gradient on irregular faces forms the
 * surface from unscaled positions.  Inputs the reference stops on (a
 * sub-patch level beyond MAX_DOMAIN = 100, (u, v) outside an irregular face,
 * the gradient at the extraordinary vertex) return UMCP_ERR_ARG naming the
 * point.  The gradient is summed per (face, slot) over the face's points in
 * the caller's order, then per vertex in the plan's fixed order: results do
 * not depend on the launch.  Needs umcp_plan; host arrays; synchronous. */
#define UMCP_POINT_MAX_BATCH (1 << 24)
int umcp_point_eval(umcp_surface *s, const double *r, int n, const int *faces, const double *uv, double *R,
                    double *nrm);
int umcp_point_grad(umcp_surface *s, const double *r, int n, const int *faces, const double *uv,
                    const double *de_dr, const double *de_dnrm, const double *frac_mult, int on_surface, double kT,
                    double *g, double *g_puv);
/* umcp_point_eval with every array on the device (d_r 3nv+3, d_faces n,
 * d_uv 2n, d_R and d_nrm 3n) and no synchronisation: the points are resolved
 * on the device by the same rules (evaluateRNRM, src/interpLib.C:7936-8196;
 * the irregular sub-patch level of src/irr_kernel.C:380-541, found with the
 * device's log10, which may differ from libm's in the last bit: a point
 * within rounding of a level border may take the adjacent sub-patch, where
 * the surface agrees to rounding), then evaluated by the same kernels, so
 * the results equal umcp_point_eval's.  A rejected point gets NaN outputs;
 * umcp_point_status (after the caller's synchronisation) then returns
 * UMCP_ERR_ARG naming the first one (*bad_point, else -1).  stream may be
 * NULL (the surface's stream).  The device calls of one surface share its
 * work space (grown on demand, the only allocation): order them on one
 * stream; calls that overlap race.  The sorts inside use memset nodes, so
 * these calls are not meant for graph capture. */
int umcp_point_eval_dev(umcp_surface *s, const double *d_r, int n, const int *d_faces, const double *d_uv,
                        double *d_R, double *d_nrm, void *stream);
/* umcp_point_grad with every array on the device (d_de_dr, d_de_dnrm 3n,
 * d_frac_mult n or NULL; d_g 3nv+3 and d_g_puv 2n accumulate) and no
 * synchronisation: resolve as umcp_point_eval_dev, a stable sort by face,
 * the faces' segments, then the same kernels and fixed-order sums, so the
 * results equal umcp_point_grad's.  A rejected point adds nothing and is
 * named by umcp_point_status. */
int umcp_point_grad_dev(umcp_surface *s, const double *d_r, int n, const int *d_faces, const double *d_uv,
                        const double *d_de_dr, const double *d_de_dnrm, const double *d_frac_mult, int on_surface,
                        double kT, double *d_g, double *d_g_puv, void *stream);
/* umcp_point_geometry with every array on the device (d_geom [n][15]), the
 * same way: rejected points (the extraordinary vertex included, as in
 * umcp_point_geometry) get NaN records */
int umcp_point_geometry_dev(umcp_surface *s, const double *d_r, int n, const int *d_faces, const double *d_uv,
                            double *d_geom, void *stream);
int umcp_point_status(umcp_surface *s, int *bad_point);
/* Local geometry at the points, replacing surface::ru, surface::rv,
 * surface::g and surface::c (src/uv_map.C:1347-1700, 884-1106, 2361-2990):
 * geom[15 i ..] = Ru[3], Rv[3] (alpha-scaled tangents along the face's u and
 * v), g (the reference's area element: |lru x lrv| of unscaled positions, on
 * irregular faces with the sub-patch derivatives scaled by the face's own u
 * and v as the reference does), c1 + c2, c1 c2, c1, c2 (principal curvatures
 * of the alpha-scaled surface, SAFETY-regularised as in src/uv_map.C:2580) and
 * the principal directions vec1[2], vec2[2] in (u, v) coordinates, unit
 * length.  The extraordinary vertex itself is rejected (the reference has no
 * sub-patch there). */
#define UMCP_POINT_GEOM 15
int umcp_point_geometry(umcp_surface *s, const double *r, int n, const int *faces, const double *uv, double *geom);
/* Control-point weights at the points, replacing surface::get_pt_coeffs and
 * get_pt_dcoeffs (src/uv_map.C:14-420): ids[18 i + x] the vertex of control
 * point x (-1 past the face's last), w[54 i + 18 d + x] its weight in the
 * position (d = 0) and in the derivatives along the face's u (1) and v (2),
 * so that R = alpha * sum w_0x (r_x + image) and likewise Ru, Rv.  At an
 * irregular face's extraordinary vertex the reference's mask: 1/2 on the
 * vertex, 1/(2 val) on its ring, derivative weights zero. */
#define UMCP_POINT_SLOTS 18
int umcp_point_coeffs(umcp_surface *s, int n, const int *faces, const double *uv, int *ids, double *w);
/* surface::penergy (src/interpLib.C:8330-8790) for n particles on the
 * limit surface, batched: particle i on face faces[i] at uv[2i], uv[2i+1],
 * with curvature preference c0[i] and area area[i]:
 *   out[i] = 1/2 kc area (c1 + c2 - c0)^2 - area * rho_f,
 * c1, c2 the principal curvatures at the point and rho_f the face's
 * curvature-energy density as faceEnergy forms it (regular faces
 * 1/2 kc (c1 + c2 - c0_f)^2, irregular faces the point energy density, over
 * dudv g0 w; src/interpLib.C:3612, 3850).  faceEnergy of a face carrying
 * particles adds the sum of its particles' out[i] (src/interpLib.C:3636,
 * 3878).  The Monge-gauge form (do_monge) is not provided. */
int umcp_particle_energy(umcp_surface *s, const double *r, int n, const int *faces, const double *uv,
                         const double *c0, const double *area, double *out);
/* The particle terms of surface::grad for n particles on the limit surface
 * (particle i on face faces[i] at uv[2i], uv[2i+1], curvature preference
 * c0[i], area area[i]), adding into g (3nv+3, with dE/dalpha) and g_puv (2n):
 * surface::pgrad (src/grad.C:3487-5336) -- the gradient of
 * 1/2 kc area (c1 + c2 - c0)^2 at each particle in the vertices, alpha and the
 * particle's (u, v) -- and the term grad and igrad add on every face
 * carrying particles whose c0 differs from the face's by 1e-7 or more
 * (src/grad.C:445-446, 458-459, 6819-6820): per Gauss point dE/d(c1 + c2) =
 * -P dAf kc (c1 + c2 - c0_face) / A with P their summed area, dAf = g0 w and
 * A = sum g w of the current surface.  With umcp_grad this gives the
 * reference's grad(r, g, puv, pg) of a surface carrying particles.  The
 * extraordinary vertex itself and points outside an irregular face's
 * sub-patches return UMCP_ERR_ARG naming the particle.  Needs umcp_plan and
 * umcp_setg0; host arrays; synchronous. */
int umcp_particle_grad(umcp_surface *s, const double *r, int n, const int *faces, const double *uv,
                       const double *c0, const double *area, double *g, double *g_puv);
/* One step of surface::nextFace (src/uv_map.C:1695-1961) for each point:
 * moves (u, v) by (du, dv) within the face, or to the border it reaches
 * first and onto the face across it (faces[i], uv updated), leaving the rest
 * of the displacement transformed so that du r_u + dv r_v is kept against the
 * old face's tangents (duv updated; 0 once inside).  mom (2 per point, may be
 * NULL) is transformed with the same matrix and, when given, coord_transform
 * (4 per point, row-major) is multiplied by it, as the reference does.  Call
 * again while any displacement remains.  The reference's random nudge for a
 * point moving along the edge it sits on (rand(), src/uv_map.C:1711-1718) is
 * not made; its deterministic 1e-12 nudges are. */
int umcp_next_face(umcp_surface *s, const double *r, int n, int *faces, double *uv, double *duv, double *mom,
                   double *coord_transform);

/* Brownian steps drawing their noise on the device: xi of step k is the
 * counter-based stream of (seed, k) that umcp_noise writes (Philox4x64-10
 * keyed by seed, Box-Muller), so a trajectory depends only on (seed, first
 * step), not on the launch or on other surfaces.  The step counter starts at
 * the umcp_set_step value (0 after umcp_plan) and advances by one per step;
 * d_energies[k] (device, capacity energies_cap, may be NULL) receives the
 * energy at the start of step k.  Asynchronous like the other _dev calls.
 * (hd.C draws with gsl_ran_gaussian, src/hd.C:1658; no bitwise counterpart.) */
int umcp_set_step(umcp_surface *s, int step);
int umcp_get_step(umcp_surface *s, int *step);
int umcp_bd_steps_rng_dev(umcp_surface *s, double *d_r, int nsteps, const umcp_bd_params *bd,
                          unsigned long long seed, double *d_energies, int energies_cap,
                          void *stream);
/* the deviates of (seed, step) for vertices [0, n): d_out[3v + k] (device) */
int umcp_noise(unsigned long long seed, int step, int n, double *d_out, void *stream);

/* Ensembles (no reference counterpart: the reference runs one surface per
 * process, src/hd.C): many independent resident surfaces stepped together.
 * Each member keeps its own tables, parameters and restraints; one step of
 * all members is one batched launch each of the restraint, face and vertex
 * kernels over device work tables (optionally captured into a hipGraph of
 * graph_steps steps and replayed).  A member's trajectory is the same alone
 * (umcp_bd_steps_rng_dev) and in any mix.  Members must be planned and
 * setg0'd, share one precision, and must not be changed while members
 * (umcp_ensemble_refresh after a change); d_r (3nv+3) and d_energies are
 * device arrays owned by the caller.  Before removing a member or freeing its
 * arrays, the caller synchronises the stream the steps ran on. */
typedef struct umcp_ensemble umcp_ensemble;
int umcp_ensemble_create(umcp_ensemble **out);
void umcp_ensemble_destroy(umcp_ensemble *e);
const char *umcp_ensemble_last_error(const umcp_ensemble *e);
int umcp_ensemble_add(umcp_ensemble *e, umcp_surface *s, double *d_r, const umcp_bd_params *bd,
                      unsigned long long seed, double *d_energies, int energies_cap);
int umcp_ensemble_remove(umcp_ensemble *e, umcp_surface *s);
int umcp_ensemble_size(const umcp_ensemble *e);
int umcp_ensemble_refresh(umcp_ensemble *e);
/* batched launch shape: point_split and flags as umcp_set_launch; face_grid
 * = 0 launches one workgroup per 64-face item, > 0 a persistent grid of that
 * many workgroups striding over the items; order 0 = member after member,
 * 1 = round-robin over members */
int umcp_ensemble_set_launch(umcp_ensemble *e, int point_split, int flags, int face_grid, int order);
int umcp_ensemble_get_launch(const umcp_ensemble *e, int *point_split, int *flags, int *face_grid,
                             int *order);
/* nsteps Brownian steps of every member with device noise (as
 * umcp_bd_steps_rng_dev); graph_steps > 0 replays a captured graph of that
 * many steps; stream NULL = the ensemble's own stream; asynchronous */
int umcp_ensemble_bd_steps(umcp_ensemble *e, int nsteps, int graph_steps, void *stream);
/* umcp_nan_check over the members: waits for the device; if some member had
 * a non-finite energy, sets *member (may be NULL) to the first such member's
 * index, clears every member's report and returns UMCP_ERR_NAN_ENERGY */
int umcp_ensemble_nan_check(umcp_ensemble *e, int *member);
/* work of the current tables (after the first umcp_ensemble_bd_steps): face
 * items (64-face groups), gather items (256-vertex blocks), members with
 * restraints */
int umcp_ensemble_work(const umcp_ensemble *e, int *face_items, int *gather_items, int *restrained);

/* Serving loop (no reference counterpart: the reference runs one simulation
 * per process).  Requests (mesh file, parameters, spontaneous curvature, N
 * Brownian steps with device noise) are queued as they arrive; preparation
 * threads admit the queue head once a slot is free and its device footprint
 * fits the budget (template parsed once per mesh file, then clone, plan, setg0,
 * start positions); the worker thread steps every admitted request together
 * (batched launches of an ensemble) in chunks of at most chunk_steps, ending
 * each chunk at the fewest remaining steps, and retires finished requests.
 * Each request's result equals running it alone (umcp_bd_steps_rng_dev). */
typedef struct umcp_server umcp_server;
typedef struct {
    double budget_bytes;   /* device bytes admitted requests may hold (<= 0: 200e9) */
    int max_active;        /* requests stepped together */
    int chunk_steps;       /* most steps between two scheduling points */
    int precision;         /* UMCP_PREC_* of every request */
    int point_split;       /* ensemble launch shape, as umcp_ensemble_set_launch */
    int flags;
    int face_grid;
    int order;
    int graph_steps;       /* > 0: replay captured graphs of up to that many steps */
    int prep_threads;      /* requests prepared concurrently (<= 0: 1) */
} umcp_server_config;
typedef struct {
    const char *mesh_path; /* .mesh file of the request's surface */
    int nv;                /* its vertex count (r0, r hold 3nv+3 doubles) */
    umcp_params params;    /* precision is the server's */
    double c0;             /* spontaneous curvature of every face (UMCP_MODE_HELFRICH) */
    int steps;
    unsigned long long seed;
    umcp_bd_params bd;
    const double *r0;      /* setg0 state (copied at submit) */
    const double *r;       /* start positions (copied at submit) */
} umcp_request;
int umcp_server_create(const umcp_server_config *cfg, umcp_server **out);
/* thread safe; *id identifies the request */
int umcp_server_submit(umcp_server *sv, const umcp_request *rq, int *id);
/* blocks until request id is done, then releases it: r_out (3nv+3), energies
 * (steps: the energy at the start of each step) and times[4] = submitted,
 * admitted, first step, done (seconds on the server's clock); each may be
 * NULL.  Returns the request's status. */
int umcp_server_wait(umcp_server *sv, int id, double *r_out, double *energies, double *times);
/* blocks until every submitted request is done */
int umcp_server_drain(umcp_server *sv);
/* writes the first min(n, 12) of: requests done, chunks, request-steps,
 * face-steps, seconds the worker spent in chunks (launch to completion),
 * peak reserved bytes, peak active requests, seconds since creation, seconds
 * rebuilding the batch's work tables, seconds retiring finished requests,
 * seconds the worker waited for work, seconds of request set-up summed over
 * the preparation threads; returns the count written */
int umcp_server_stats(umcp_server *sv, double *stats, int n);
const char *umcp_server_last_error(const umcp_server *sv);
/* finishes admitted requests, drops queued ones, joins the threads */
void umcp_server_destroy(umcp_server *sv);

/* launch shape of the face kernel (no reference counterpart; set by the
 * tuner): point_split in 1..4 waves share the Gauss points of 64 faces
 * (their control points staged in LDS); flags: UMCP_LAUNCH_OCC2 bounds
 * registers (two waves per SIMD in FP64, four in the mixed VALU path, two on
 * the matrix-core path); UMCP_LAUNCH_MFMA (mixed precision only, ignored in
 * FP64) runs the basis contractions on the matrix cores, with point_split 1,
 * 2 or 4; UMCP_LAUNCH_CPS (point_split 2, 3 or 4, not with MFMA): the waves
 * of a group also split the control points, each accumulating the gradient
 * of its own share over every Gauss point (the points' reverse passes
 * through LDS), which needs about half the registers */
#define UMCP_LAUNCH_OCC2 2
#define UMCP_LAUNCH_MFMA 4
#define UMCP_LAUNCH_CPS 8
int umcp_set_launch(umcp_surface *s, int point_split, int flags);
int umcp_get_launch(const umcp_surface *s, int *point_split, int *flags);

/* per-launch timing with HIP events recorded on the stream each launch uses
 * (no reference counterpart; benchmark support): begin, run work, then end
 * returns the summed milliseconds and the number of launches per entry
 * (max_cls entries): 0 the face kernel (all face classes share one launch),
 * 1 the vertex gather (gradient or Brownian update), 2 the restraint
 * pre-pass and finalize (surfaces with global restraints) */
int umcp_profile_begin(umcp_surface *s);
/* while profiling, time one evaluation's face launch in every `every`
 * (default 1), so the events' own cost stays off the other steps */
int umcp_profile_sample(umcp_surface *s, int every);
int umcp_profile_end(umcp_surface *s, int max_cls, double *ms, int *launches);

/* introspection (tests): the shared table of a face class and its points */
int umcp_class_count(const umcp_surface *s);
int umcp_class_info(const umcp_surface *s, int cls, int *valence, int *ncp, int *npt, int *nf);
int umcp_class_table(const umcp_surface *s, int cls, double *table, double *weights);
int umcp_face_points(const umcp_surface *s, int face, int *ncp, int *cp);
int umcp_face_lipids(const umcp_surface *s, int face, double *Lo, double *Li);
/* each lipid type's area in the leaflets of face's triangle after setg0
 * (theTriangles[t].composition.outerLeaflet / innerLeaflet): Lo, Li hold
 * umcp_lipid_type_count values */
int umcp_face_lipid_types(const umcp_surface *s, int face, double *Lo, double *Li);

#ifdef __cplusplus
}
#endif
#endif
