/*
 * umcp_oracle.h -- CPU oracle for the membrane-surface hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C, FP64, serial restatement of
 * the reference (alexsodt/UMCP) algorithm: mesh load + orientation
 * (src/interpLib.C:1144-2606), plan / formula tables (src/interpLib.C:
 * 4998-6670, src/gauss.C:27-299), reference state (src/interpLib.C:2969-3388),
 * energy (src/interpLib.C:3405-3881, 4180-4352), gradient (src/grad.C:65-942,
 * 6445-7457) and the per-vertex Brownian step (docs/timescales.tex:59-67,
 * src/bd.C:1035-1036).  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker.  The product library
 * (umcp_amd/libumcp_core.so) never links or calls it.
 */
#ifndef UMCP_ORACLE_H
#define UMCP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#include <stdint.h>

typedef struct omesh omesh;

/* energy modes: 0 = reference default (FIXED_A + LOCAL_LIPID_ENERGY,
 * src/local_config.h:21-22); 1 = per-face c0 Helfrich with g0 area element
 * (the #else branches of src/interpLib.C:3766-3844).
 * grad_form (mode 1 only): 0 = surface::grad's gradient, whose regular faces
 * add 1/2 kc (c1+c2-c0)^2 dudv w to dE/dg (src/grad.C:448; igrad does not,
 * src/grad.C:6822-6824); 1 = the exact derivative of the energy. */
typedef struct {
    double kc, kg, KA, k_reg;
    int mode;
    double lipid_c0, lipid_apl;   /* single lipid type, default "ZERO": 0, 68.3 */
    int grad_form;
} oparams;

int    orc_load(const char *path, omesh **out);        /* loadLattice + orientSurface */
void   orc_free(omesh *m);
int    orc_nv(const omesh *m);
int    orc_nfaces(const omesh *m, int *nreg, int *nirr);
void   orc_get_positions(const omesh *m, double *r);   /* r[3nv+3], alpha = 1 */
int    orc_plan(omesh *m);                              /* generatePlan */
void   orc_set_params(omesh *m, const oparams *p);
void   orc_set_face_c0(omesh *m, const double *c0);    /* per face, face order */
void   orc_setg0(omesh *m, const double *r);            /* setg0 */
double orc_energy(omesh *m, const double *r);           /* surface::energy */
double orc_face_energy(omesh *m, int f, const double *r);
void   orc_grad(omesh *m, const double *r, double *g);  /* grad + igrad, accumulates */
void   orc_area(omesh *m, const double *r, double *A, double *A0);
void   orc_bd_steps(omesh *m, double *r, const double *xi, int nsteps,
                    double mobility_dt, double noise_amp, double *energies);

/* global restraints (area: src/interpLib.C:4322-4326; volume: src/minimize.C:
 * 90-114 over src/interpLib.C:997-1098).  A0 <= 0: sum(dudv g0 w) of setg0;
 * V0 == 0: the volume at setg0.  orc_energy and orc_grad include them. */
void   orc_set_restraints(omesh *m, double KA_global, double KV, double A0, double V0);
void   orc_limit_points(const omesh *m, const double *r, double *L);   /* 3 nv */
double orc_volume(const omesh *m, const double *r);
void   orc_volume_grad(const omesh *m, const double *r, double *dV);  /* 3 nv */
double orc_restraint_energy(omesh *m, const double *r, double *lamA, double *lamV);

/* sum over the Gauss points of the regular faces of
 * 1/2 kc (c1 + c2 - c0)^2 dudv w g(r), with c1 + c2 evaluated at r_star and
 * g(r) the area element at r (Helfrich form): the function whose gradient in
 * r at r = r_star is what surface::grad's dE/dg term of src/grad.C:448 adds
 * to the exact gradient */
double orc_frozen_curvature_area(omesh *m, const double *r_star, const double *r);

/* effective-mass integrators of src/hd.C (Cartesian momenta): the point
 * mass of the ppf lattice points per face, the dense mass matrix M (nv x nv),
 * its inverse by Cholesky (0 or failing pivot + 1), the reference's
 * truncation of the inverse, and nsteps of the integrator (mode 0 =
 * do_bd_membrane, 1 = do_ld with OLD_LANGEVIN friction); pp[3nv] momenta,
 * xi[nsteps][3nv] standard normals, dt in s, kT in kcal/mol, gamma in the
 * reference's units (gamma_langevin) */
int    orc_effm_points_per_face(void);
double orc_effm_point_mass(omesh *m);
void   orc_effm_mass(omesh *m, double mass_per_pt, double *M);
int    orc_effm_invert(int n, double *A);
long   orc_effm_truncate(int n, double *Mi, double rel_tol);
void   orc_effm_steps(omesh *m, double *r, double *pp, const double *xi, int nsteps, int mode, double dt,
                      double kT, double gamma, const double *Mi, double *energies);

/* energy + gradient with the faces split over nthreads OpenMP threads
 * (per-thread gradient buffers summed in thread order); g overwritten */
double orc_energy_grad_threads(omesh *m, const double *r, double *g, int nthreads);
void   orc_bd_steps_threads(omesh *m, double *r, const double *xi, int nsteps, double mobility_dt,
                            double noise_amp, double *energies, int nthreads);

/* introspection for tests */
int    orc_face_info(const omesh *m, int f, int *owner, int *ncp, int *cp, int *tri);
int    orc_face_table(const omesh *m, int f, int p, double *w6, double *weight,
                      double *pbc, double *g0);
void   orc_tri_lipids(const omesh *m, int t, double *Lo, double *Li);
/* lipid compositions (src/lipid_composition.C:47-333, src/interpLib.C:
 * 3127-3196, 3286-3358): n types with spontaneous curvature c0, area per
 * lipid apl and input parts per leaflet, packed at the next setg0 with
 * rand() after srand(seed); -1 for a bad composition */
#define ORC_MAX_LIPID_TYPES 64
int    orc_set_lipids(omesh *m, int n, const double *c0, const double *apl, const double *in_parts,
                      const double *out_parts, unsigned int seed);
void   orc_tri_lipid_types(const omesh *m, int t, double *Lo, double *Li);
/* glibc's rand(), restated (srand(seed), then rand()) */
typedef struct { int32_t r[31]; int f, b; } orc_grand;
void   orc_grand_seed(orc_grand *g, unsigned int seed);
int    orc_grand_next(orc_grand *g);
int    orc_ntri(const omesh *m);
int    orc_tri_order_face(const omesh *m, int t);
void   orc_neighbors(const omesh *m, int v, int *val, int *nb);

/* quadrature restatement (src/gauss.C:96-299 over QUADPACK QAGS) */
void   orc_get_v_pts(double *v_pts, double *w_pts, int npts, double k);
void   orc_get_u_pts(double *u_pts, double *w_pts, int npts, double k, double v);
double orc_qags(double (*f)(double, void *), void *ctx, double a, double b,
                double epsabs, double epsrel, int limit, double *abserr, int *ier, int *last);

/* basis restatement: the 12 quartic box splines and derivatives at (u,v),
 * in point_list order (src/interpLib.C:5836-5918, 6183-6189) */
void   orc_basis(double u, double v, double out[6][12]);
/* irregular (Stam) table for valence val at parameter (u,v): out[6][val+6] */
int    orc_irregular_table(int val, double u, double v, double *out);

/* points on the limit surface (face in the reference's numbering: regular
 * faces first).  orc_point_eval: surface::evaluateRNRM, position (alpha
 * scaled) and unit normal; at an irregular face's extraordinary vertex
 * (u = v = 0) the reference's limit mask and a NaN normal.  orc_point_grad:
 * surface::pointGradient, adds to gr[3nv+3] and g_puv[2].  Both return -1 for
 * the inputs the reference rejects with exit(1) (or, for the gradient at the
 * extraordinary vertex, dereferences a NULL map). */
int    orc_point_eval(const omesh *m, const double *r, int f, double u, double v, double *rp,
                      double *nrm);
int    orc_point_grad(const omesh *m, const double *r, int f, double u, double v,
                      const double *de_dr, const double *de_dnrm, double frac_mult, int on_surface,
                      double kT, double *gr, double *g_puv);

/* point geometry (surface::ru, rv, g, c): out[15] = Ru[3], Rv[3], g,
 * c1 + c2, c1 c2, c1, c2, principal directions vec1[2], vec2[2] in (u, v) */
int    orc_point_geometry(const omesh *m, const double *r, int f, double u, double v, double *out);

/* Particle terms of the gradient: surface::pgrad (src/grad.C:3487-5336) for
 * n particles (face, (u, v), c0, area) -- E_i = 1/2 kc a_i (c1 + c2 - c0_i)^2
 * at the particle; dE/dr and dE/dalpha added into g (3nv+3), dE/du, dE/dv into
 * g_puv (2n) -- plus the terms grad and igrad add on faces carrying particle
 * area (src/grad.C:445-446, 458-459, 6819-6820).  Returns 0, or -1-i for a
 * particle i the reference would stop on. */
int    orc_particle_grad(omesh *m, const double *r, int n, const int *faces, const double *uv, const double *c0,
                         const double *area, double *g, double *g_puv);
/* pins: the particles' curvature energy (whose gradient pgrad is), and the
 * frozen-area density function whose gradient at r_star is grad's term */
double orc_particle_curvature_energy(const omesh *m, const double *r, int n, const int *faces, const double *uv,
                                     const double *c0, const double *area);
double orc_particle_density_frozen(omesh *m, const double *r_star, const double *r, int n, const int *faces,
                                   const double *c0, const double *area);
void   orc_basis10(double u, double v, double out[10][12]);
int    orc_face_density(omesh *m, int f, const double *r, double *num, double *area);
int    orc_particle_energy(const omesh *m, const double *r, int f, double u, double v, double pc0, double pa,
                           double fed, double *out);

#ifdef __cplusplus
}
#endif
#endif
