// The device-resident surface object and the extern "C" boundary (umcp.h).
//
// Lifecycle mirrors the reference driver (src/min.C:68-99, src/hd.C:162-273):
// loadLattice -> readLipidComposition (default single "ZERO" lipid) ->
// generatePlan -> setg0 -> energy / grad per step.  The mesh, the shared
// Gauss-point tables, the per-face reference state and all work buffers stay
// in HBM between calls; host arrays cross the boundary only in the host-array
// entry points.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <set>
#include <string>
#include <utility>
#include <vector>

#include "../../include/umcp.h"
#include "surface_host.hpp"
#include "umcp_kernels.hpp"
#include "umcp_state.hpp"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

using namespace umcp;

namespace {

thread_local int g_device = 0;

}  // namespace

namespace {

int fail(umcp_surface *s, int code, const std::string &msg)
{
    if (s) s->err = msg;
    return code;
}

int hip_fail(umcp_surface *s, hipError_t e, const char *where)
{
    return fail(s, UMCP_ERR_DEVICE, std::string(where) + ": " + hipGetErrorString(e));
}

// device arrays of the plan are carved from the surface's arena (one
// hipMalloc of plan_bytes, 256-byte aligned slices); anything beyond it gets
// an allocation of its own
template <typename T>
int dev_alloc(umcp_surface *s, T **p, size_t n)
{
    if (n == 0) n = 1;
    const size_t bytes = n * sizeof(T);
    const size_t off = (s->arena_used + kArenaAlign - 1) & ~(kArenaAlign - 1);
    if (s->arena && off + bytes <= s->arena_size) {
        s->arena_used = off + bytes;
        *p = (T *)(s->arena + off);
        return UMCP_OK;
    }
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) return fail(s, UMCP_ERR_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
    s->allocations.push_back(q);
    s->device_bytes += bytes;
    *p = (T *)q;
    return UMCP_OK;
}

// the Gauss-point tables of a face kind are the same for every surface: one
// upload per device and kind
int upload_tables_once(umcp_surface *s, const FaceClass &C)
{
    static std::mutex mu;
    static std::set<std::pair<int, int>> done;
    const int kind = kind_of(C.ncp, C.npt);
    std::lock_guard<std::mutex> lk(mu);
    if (done.count({s->device, kind})) return UMCP_OK;
    hipError_t e = upload_tables(kind, C.table.data(), C.weight.data(), C.corner.data(), C.npt, C.ncp);
    if (e != hipSuccess) return hip_fail(s, e, "upload Gauss-point tables");
    done.insert({s->device, kind});
    return UMCP_OK;
}

template <typename T>
int dev_upload(umcp_surface *s, T *dst, const T *src, size_t n)
{
    if (n == 0) return UMCP_OK;
    hipError_t e = hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, s->stream);
    if (e != hipSuccess) return hip_fail(s, e, "upload");
    return UMCP_OK;
}

int sync(umcp_surface *s)
{
    hipError_t e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) return hip_fail(s, e, "hipStreamSynchronize");
    return UMCP_OK;
}

PhysParams phys(const umcp_surface *s)
{
    PhysParams ph;
    ph.kc = s->prm.kc;
    ph.kg = s->prm.kg;
    ph.KA = s->prm.KA;
    ph.k_reg = s->prm.k_reg;
    ph.lipid_c0 = s->prm.lipid_c0;
    ph.mode = s->prm.mode;
    ph.grad_exact = s->prm.grad_form == UMCP_GRAD_EXACT;
    return ph;
}

}  // namespace

namespace umcp {

bool restraints_on(const umcp_surface *s) { return s->rsp.KA_global != 0.0 || s->rsp.KV != 0.0; }

FaceArgs face_args(umcp_surface *s, size_t c, const double *d_r)
{
    const FaceClass &C = s->P.classes[c];
    const ClassDev &D = s->cls[c];
    FaceArgs a{};
    a.r = d_r;
    a.nv = s->T.nv;
    a.nf = C.nf();
    a.cp = D.cp;
    a.img = D.img;
    a.has_img = D.has_img;
    a.g0 = D.g0;
    a.rurv0 = D.rurv0;
    a.ig0 = D.ig0;
    a.lw = D.lw;
    a.ls = D.ls;
    a.lv = D.lv;
    a.c0 = D.c0;
    a.dudv = C.dudv;
    std::memcpy(a.pbc, s->T.pbc, sizeof a.pbc);
    a.gbuf = s->d_gbuf;
    a.gbuf32 = s->d_gbuf32;
    a.gl_start = s->d_gl_start;
    a.linc_start = s->d_linc_start;
    a.linc = s->d_linc;
    a.lvert = s->d_lvert;
    a.loc = D.loc;
    a.ph = phys(s);
    a.step_ctr = nullptr;
    a.partial = s->d_partial;
    a.rs = restraints_on(s) ? s->d_rs : nullptr;
    a.rpart = s->d_rpart;
    a.ngroups = s->npartial;
    a.rs_fold = a.rs && a.ngroups <= kFoldMaxGroups ? 1 : 0;
    a.KA_g = s->rsp.KA_global;
    a.KV = s->rsp.KV;
    a.A0 = s->rsp.A0 > 0.0 ? s->rsp.A0 : s->A0_setg0;
    a.V0 = s->rsp.V0 != 0.0 ? s->rsp.V0 : s->V_setg0;
    a.do_area = s->rsp.KA_global != 0.0 ? 1 : 0;
    a.do_volume = s->rsp.KV != 0.0 ? 1 : 0;
    return a;
}

// mirrors every dev_alloc of umcp_plan (zero-sized requests allocate one element)
size_t plan_bytes(const umcp_surface *s)
{
    auto sz = [](size_t n, size_t elem) { return ((n ? n : 1) * elem + kArenaAlign - 1) & ~(kArenaAlign - 1); };
    size_t b = 0;
    int metric_total = 0, ngroups = 0;
    for (const FaceClass &C : s->P.classes) {
        const size_t nf = C.nf();
        b += sz(C.cp.size(), 4) + sz(C.pos.size(), 4) + sz(C.loc.size(), 2) + sz(C.img.size(), 1) + sz(nf, 1);
        b += 3 * sz(nf * C.npt, 8) + 3 * sz(nf, 8);
        metric_total += metric_blocks(C.nf());
        ngroups += face_groups(C.nf());
    }
    const size_t nv = s->T.nv;
    b += 2 * sz(3 * nv + 3, 8) + sz(3 * (size_t)s->P.nnz, 8) + sz(3 * (size_t)s->P.nnz, 4);
    size_t nloc = 0;
    for (const FaceClass &C : s->P.classes) nloc += C.gl_vert.size();
    b += sz(4 * (size_t)std::max(ngroups, metric_total), 8) + sz(8, 8) + sz(s->P.vstart.size(), 4);
    // group records: per-group bases, incidence ranges and lists, vertex CSR
    b += sz(ngroups + 1, 4) + sz(nloc + 1, 4) + sz(s->P.nnz, 2) + sz(nv + 1, 4) + 2 * sz(nloc, 4);
    b += sz(8, 8) + sz(2 * (size_t)ngroups, 8) + 3 * sz(1, 4);   // rs, rpart, step, fin, nan
    b += 2 * sz(s->P.classes.size(), sizeof(FaceArgs)) + sz(ngroups, sizeof(FaceItem));
    return b;
}

}  // namespace umcp

namespace {

// volume of the limit-point polyhedron of d_r through the restraint pre-pass
// (a volume-only face table); the result lands in d_rs[4]
int enqueue_volume(umcp_surface *s, const double *d_r, hipStream_t st)
{
    std::vector<FaceArgs> fa(s->P.classes.size());
    for (size_t c = 0; c < fa.size(); c++) {
        fa[c] = face_args(s, c, d_r);
        fa[c].rs = s->d_rs;
        fa[c].rs_fold = 0;   // no face kernel follows: the finalize writes d_rs
        fa[c].do_area = 0;
        fa[c].do_volume = 1;
        fa[c].KA_g = 0.0;
        fa[c].KV = 0.0;
    }
    hipError_t e = hipMemcpyAsync(s->d_fa_aux, fa.data(), fa.size() * sizeof(FaceArgs), hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return hip_fail(s, e, "upload face table");
    BatchFaceArgs fb{s->d_fa_aux, s->d_items, s->npartial, d_r, nullptr, s->d_fin, 1};
    e = launch_restraint_prepass(fb, st);
    if (e != hipSuccess) return hip_fail(s, e, "restraint kernel launch");
    return UMCP_OK;
}

// irregular classes first so their few blocks start early and overlap the
// regular faces
void layout_launch(umcp_surface *s)
{
    s->launch_order.clear();
    for (size_t c = 1; c < s->cls.size(); c++)
        if (s->P.classes[c].nf() > 0) s->launch_order.push_back((int)c);
    if (!s->P.classes.empty() && s->P.classes[0].nf() > 0) s->launch_order.push_back(0);
    s->block_start.assign(1, 0);
    for (int c : s->launch_order) s->block_start.push_back(s->block_start.back() + face_groups(s->P.classes[c].nf()));
    s->npartial = s->block_start.back();
}

// enqueue one energy (+ gradient) evaluation on the surface stream
int prof_pair(umcp_surface *s, int cls, hipEvent_t *ev)
{
    if (s->prof_used * 2 + 2 > s->prof_events.size()) {
        for (int k = 0; k < 2; k++) {
            hipEvent_t e;
            hipError_t err = hipEventCreate(&e);
            if (err != hipSuccess) return hip_fail(s, err, "hipEventCreate");
            s->prof_events.push_back(e);
        }
        s->prof_class.push_back(cls);
    }
    s->prof_class[s->prof_used] = cls;
    ev[0] = s->prof_events[2 * s->prof_used];
    ev[1] = s->prof_events[2 * s->prof_used + 1];
    s->prof_used++;
    return UMCP_OK;
}

// the per-class face tables on the device, refreshed when parameters changed
int sync_face_table(umcp_surface *s, hipStream_t st)
{
    if (!s->fa_dirty) return UMCP_OK;
    s->h_fa.resize(s->P.classes.size());
    for (size_t c = 0; c < s->P.classes.size(); c++) s->h_fa[c] = face_args(s, c, s->d_r);
    hipError_t e = hipMemcpyAsync(s->d_fa_tab, s->h_fa.data(), s->h_fa.size() * sizeof(FaceArgs),
                                  hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return hip_fail(s, e, "upload face table");
    s->fa_dirty = false;
    return UMCP_OK;
}

int enqueue_eval(umcp_surface *s, const double *d_r, hipStream_t st, GatherArgs ga, int *step_ctr = nullptr)
{
    const bool rson = restraints_on(s);
    if (int rc = sync_face_table(s, st)) return rc;
    BatchFaceArgs fb{};
    fb.fa = s->d_fa_tab;
    fb.items = s->d_items;
    fb.nitems = s->npartial;
    fb.r = d_r;
    fb.step_ctr = step_ctr;
    fb.fin = s->d_fin;
    fb.nfin = rson && s->npartial > kFoldMaxGroups ? 1 : 0;   // else folded into the face kernel
    hipEvent_t ev[2] = {nullptr, nullptr};
    const bool timed = s->prof_on && (s->prof_tick++ % s->prof_every) == 0;
    if (rson) {
        // timed evaluations time the restraint pre-pass and finalize too (entry 2)
        hipEvent_t evr[2] = {nullptr, nullptr};
        if (timed) {
            if (int rc = prof_pair(s, 2, evr)) return rc;
            (void)hipEventRecord(evr[0], st);
        }
        hipError_t e = launch_restraint_prepass(fb, st);
        if (e != hipSuccess) return hip_fail(s, e, "restraint kernel launch");
        if (timed) (void)hipEventRecord(evr[1], st);
    }
    if (timed) {
        int rc = prof_pair(s, 0, ev);
        if (rc) return rc;
        (void)hipEventRecord(ev[0], st);
    }
    const bool mixed = s->prm.precision == UMCP_PREC_MIXED;
    hipError_t e = launch_faces_batch(fb, 0, s->point_split, s->launch_flags, mixed, features_of(phys(s)), st);
    if (e != hipSuccess) return hip_fail(s, e, "face kernel launch");
    if (timed) (void)hipEventRecord(ev[1], st);
    ga.nv = s->T.nv;
    ga.vstart = s->d_lvstart;
    ga.vrec = s->d_lvrec;
    ga.nan_flag = s->d_nan;
    ga.gbuf = s->d_gbuf;
    ga.gbuf32 = mixed ? s->d_gbuf32 : nullptr;
    ga.partial = s->d_partial;
    ga.npartial = s->npartial;
    if (rson) ga.rs = s->d_rs;
    hipEvent_t evg[2] = {nullptr, nullptr};
    if (timed) {   // the vertex gather (entry 1)
        if (int rc = prof_pair(s, 1, evg)) return rc;
        (void)hipEventRecord(evg[0], st);
    }
    e = launch_vertex_gather(ga, st);
    if (e != hipSuccess) return hip_fail(s, e, "gather kernel launch");
    if (timed) (void)hipEventRecord(evg[1], st);
    return UMCP_OK;
}

// UMCP_ERR_NAN_ENERGY if an evaluation since the last report had a
// non-finite energy (the flag is then cleared).  device_sync: wait for the
// whole device (device entry points may run on the caller's streams), else
// for the surface's stream.
int nan_status(umcp_surface *s, bool device_sync)
{
    hipError_t e = device_sync ? hipDeviceSynchronize() : hipStreamSynchronize(s->stream);
    if (e != hipSuccess) return hip_fail(s, e, "synchronize");
    int flag = 0;
    e = hipMemcpyAsync(&flag, s->d_nan, sizeof flag, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) return hip_fail(s, e, "download nan flag");
    if (!flag) return UMCP_OK;
    const int zero = 0;
    e = hipMemcpyAsync(s->d_nan, &zero, sizeof zero, hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) return hip_fail(s, e, "clear nan flag");
    return fail(s, UMCP_ERR_NAN_ENERGY, "nan energy");
}

int check_ready(umcp_surface *s, bool need_g0 = true)
{
    if (!s) return UMCP_ERR_ARG;
    if (!s->planned) return fail(s, UMCP_ERR_STATE, "plan has not been generated (call umcp_plan)");
    if (need_g0 && !s->g0_set) return fail(s, UMCP_ERR_STATE, "reference state not set (call umcp_setg0)");
    hipError_t e = hipSetDevice(s->device);
    if (e != hipSuccess) return hip_fail(s, e, "hipSetDevice");
    return UMCP_OK;
}

int upload_r(umcp_surface *s, const double *r)
{
    return dev_upload(s, s->d_r, r, 3 * (size_t)s->T.nv + 3);
}

}  // namespace

extern "C" {

int umcp_link_check(void) { return 1; }

int umcp_set_device(int device)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return UMCP_ERR_DEVICE;
    g_device = device;
    return UMCP_OK;
}

int umcp_load_mesh(const char *path, umcp_surface **out)
{
    if (!path || !out) return UMCP_ERR_ARG;
    *out = nullptr;
    auto *s = new umcp_surface();
    s->device = g_device;
    std::string err;
    int rc = parse_mesh(path, s->T, err);
    if (rc == 0) rc = orient_mesh(s->T, err);
    if (rc != 0) {
        s->err = err;
        *out = s;  // caller may read umcp_last_error, then destroy
        return UMCP_ERR_BAD_MESH;
    }
    // default parameters: reference globals (src/interpLib.C:85-95) and the
    // "ZERO" lipid (src/lipid_composition.C:18-22)
    s->prm.kc = 14.0;
    s->prm.kg = 0.0;
    s->prm.KA = 0.0;
    s->prm.k_reg = 0.0;
    s->prm.mode = UMCP_MODE_LIPID;
    s->prm.precision = UMCP_PREC_FP64;
    s->prm.lipid_c0 = 0.0;
    s->prm.lipid_apl = 68.3;
    s->prm.grad_form = UMCP_GRAD_REFERENCE;
    *out = s;
    return UMCP_OK;
}

void umcp_destroy(umcp_surface *s)
{
    if (!s) return;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (void *p : s->allocations) (void)hipFree(p);
    if (s->pt_stage) (void)hipHostFree(s->pt_stage);
    if (s->pt_host) (void)hipHostFree(s->pt_host);
    for (hipEvent_t e : s->prof_events) (void)hipEventDestroy(e);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

const char *umcp_last_error(const umcp_surface *s) { return s ? s->err.c_str() : "null handle"; }

int umcp_nv(const umcp_surface *s) { return s ? s->T.nv : UMCP_ERR_ARG; }

int umcp_nfaces(const umcp_surface *s, int *nreg, int *nirr)
{
    if (!s) return UMCP_ERR_ARG;
    if (!s->prepared) return 0;
    if (nreg) *nreg = s->P.nreg;
    if (nirr) *nirr = s->P.nfaces - s->P.nreg;
    return s->P.nfaces;
}

int umcp_get(const umcp_surface *s, double *r)
{
    if (!s || !r) return UMCP_ERR_ARG;
    std::memcpy(r, s->T.r.data(), sizeof(double) * 3 * (size_t)s->T.nv);
    r[3 * s->T.nv] = r[3 * s->T.nv + 1] = r[3 * s->T.nv + 2] = 1.0;
    return UMCP_OK;
}

int umcp_set_params(umcp_surface *s, const umcp_params *p)
{
    if (!s || !p) return UMCP_ERR_ARG;
    if (p->mode != UMCP_MODE_LIPID && p->mode != UMCP_MODE_HELFRICH)
        return fail(s, UMCP_ERR_ARG, "unknown energy mode");
    if (p->precision != UMCP_PREC_FP64 && p->precision != UMCP_PREC_MIXED)
        return fail(s, UMCP_ERR_ARG, "unknown precision");
    if (p->grad_form != UMCP_GRAD_REFERENCE && p->grad_form != UMCP_GRAD_EXACT)
        return fail(s, UMCP_ERR_ARG, "unknown gradient form");
    if (!(p->lipid_apl > 0)) return fail(s, UMCP_ERR_ARG, "lipid area must be positive");
    bool repack = s->g0_set && p->lipid_apl != s->prm.lipid_apl;
    s->prm = *p;
    s->fa_dirty = true;
    if (repack) s->g0_set = false;  // lipid packing depends on the lipid area
    return UMCP_OK;
}

int umcp_prepare(umcp_surface *s)
{
    if (!s) return UMCP_ERR_ARG;
    if (s->prepared) return UMCP_OK;
    std::string err;
    if (build_plan(s->T, s->P, err) != 0) return fail(s, UMCP_ERR_BAD_MESH, err);
    s->prepared = true;
    return UMCP_OK;
}

int umcp_clone(const umcp_surface *tmpl, umcp_surface **out)
{
    if (!tmpl || !out) return UMCP_ERR_ARG;
    *out = nullptr;
    if (!tmpl->prepared) return UMCP_ERR_STATE;
    auto *s = new umcp_surface();
    s->device = g_device;
    s->T = tmpl->T;
    s->P = tmpl->P;
    s->prm = tmpl->prm;
    s->prepared = true;
    *out = s;
    return UMCP_OK;
}

int umcp_device_bytes(const umcp_surface *s, size_t *planned, size_t *allocated)
{
    if (!s) return UMCP_ERR_ARG;
    if (!s->prepared) return UMCP_ERR_STATE;
    if (planned) *planned = plan_bytes(s);
    if (allocated) *allocated = s->device_bytes;
    return UMCP_OK;
}

int umcp_plan(umcp_surface *s)
{
    if (!s) return UMCP_ERR_ARG;
    if (s->planned) return UMCP_OK;
    int rc0 = umcp_prepare(s);
    if (rc0) return rc0;
    hipError_t e = hipSetDevice(s->device);
    if (e != hipSuccess) return hip_fail(s, e, "hipSetDevice");
    if (!s->stream) {
        e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
        if (e != hipSuccess) return hip_fail(s, e, "hipStreamCreate");
    }
    const int nv = s->T.nv;
    int rc;
    s->arena_size = plan_bytes(s);
    s->arena_used = 0;
    e = hipMalloc((void **)&s->arena, s->arena_size);
    if (e != hipSuccess) {
        s->arena = nullptr;
        return fail(s, UMCP_ERR_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    s->allocations.push_back(s->arena);
    s->device_bytes += s->arena_size;
    s->face_c0.assign(s->P.nfaces, 0.0);
    s->face_Lo.assign(s->P.nfaces, 0.0);
    s->face_Li.assign(s->P.nfaces, 0.0);
    s->cls.assign(s->P.classes.size(), ClassDev());
    for (size_t c = 0; c < s->P.classes.size(); c++) {
        const FaceClass &C = s->P.classes[c];
        ClassDev &D = s->cls[c];
        const size_t nf = C.nf();
        if (kind_of(C.ncp, C.npt) < 0) return fail(s, UMCP_ERR_BAD_MESH, "no kernel for this face valence");
        if ((rc = dev_alloc(s, &D.cp, C.cp.size())) || (rc = dev_alloc(s, &D.pos, C.pos.size())) || (rc = dev_alloc(s, &D.img, C.img.size())) ||
            (rc = dev_alloc(s, &D.loc, C.loc.size())) ||
            (rc = dev_alloc(s, &D.has_img, nf)) || (rc = dev_alloc(s, &D.g0, nf * C.npt)) ||
            (rc = dev_alloc(s, &D.rurv0, nf * C.npt)) || (rc = dev_alloc(s, &D.ig0, nf * C.npt)) || (rc = dev_alloc(s, &D.lw, nf)) ||
            (rc = dev_alloc(s, &D.ls, nf)) || (rc = dev_alloc(s, &D.lv, nf)) || (rc = dev_alloc(s, &D.c0, nf)) ||
            0)
            return rc;
        if ((rc = upload_tables_once(s, C))) return rc;
        if ((rc = dev_upload(s, D.cp, C.cp.data(), C.cp.size())) ||
            (rc = dev_upload(s, D.pos, C.pos.data(), C.pos.size())) ||
            (rc = dev_upload(s, D.loc, C.loc.data(), C.loc.size())) ||
            (rc = dev_upload(s, D.img, C.img.data(), C.img.size())) ||
            (rc = dev_upload(s, D.has_img, C.has_img.data(), nf)) ||
            0)
            return rc;
        std::vector<double> zero(nf * C.npt, 0.0);
        if ((rc = dev_upload(s, D.c0, zero.data(), nf)) || (rc = dev_upload(s, D.lw, zero.data(), nf)) ||
            (rc = dev_upload(s, D.ls, zero.data(), nf)) || (rc = dev_upload(s, D.lv, zero.data(), nf)) ||
            (rc = dev_upload(s, D.g0, zero.data(), nf * C.npt)) ||
            (rc = dev_upload(s, D.rurv0, zero.data(), nf * C.npt)) ||
            (rc = dev_upload(s, D.ig0, zero.data(), nf * C.npt)))
            return rc;
    }
    layout_launch(s);
    int metric_total = 0;
    for (size_t c = 0; c < s->P.classes.size(); c++) metric_total += metric_blocks(s->P.classes[c].nf());
    const int max_partial = std::max(s->npartial, metric_total);
    if ((rc = dev_alloc(s, &s->d_r, 3 * (size_t)nv + 3)) || (rc = dev_alloc(s, &s->d_g, 3 * (size_t)nv + 3)) ||
        (rc = dev_alloc(s, &s->d_gbuf, 3 * (size_t)s->P.nnz)) || (rc = dev_alloc(s, &s->d_gbuf32, 3 * (size_t)s->P.nnz)) ||
        (rc = dev_alloc(s, &s->d_partial, 4 * (size_t)max_partial)) || (rc = dev_alloc(s, &s->d_scalar, 8)) ||
        (rc = dev_alloc(s, &s->d_vstart, s->P.vstart.size())))
        return rc;
    if ((rc = dev_upload(s, s->d_vstart, s->P.vstart.data(), s->P.vstart.size()))) return rc;
    // group records in launch-slot order: slot -> first record, record ->
    // incidence range, and the vertex -> records CSR of the gather
    {
        std::vector<int> gl_start(1, 0), linc_start(1, 0), lvert;
        std::vector<uint16_t> linc;
        for (size_t i = 0; i < s->launch_order.size(); i++) {
            const FaceClass &C = s->P.classes[s->launch_order[i]];
            for (int g = 0; g + 1 < (int)C.gl_off.size(); g++) {
                for (int l = C.gl_off[g]; l < C.gl_off[g + 1]; l++) {
                    lvert.push_back(C.gl_vert[l]);
                    linc.insert(linc.end(), C.gl_inc.begin() + C.gl_inc_off[l], C.gl_inc.begin() + C.gl_inc_off[l + 1]);
                    linc_start.push_back((int)linc.size());
                }
                gl_start.push_back((int)lvert.size());
            }
        }
        std::vector<int> lvstart(nv + 1, 0), lvrec(lvert.size());
        for (int v : lvert) lvstart[v + 1]++;
        for (int v = 0; v < nv; v++) lvstart[v + 1] += lvstart[v];
        std::vector<int> fill(lvstart.begin(), lvstart.end() - 1);
        for (size_t L = 0; L < lvert.size(); L++) lvrec[fill[lvert[L]]++] = (int)L;
        s->nrecords = (int)lvert.size();
        if ((rc = dev_alloc(s, &s->d_gl_start, gl_start.size())) || (rc = dev_alloc(s, &s->d_linc_start, linc_start.size())) ||
            (rc = dev_alloc(s, &s->d_linc, linc.size())) || (rc = dev_alloc(s, &s->d_lvstart, lvstart.size())) ||
            (rc = dev_alloc(s, &s->d_lvrec, lvrec.size())) || (rc = dev_alloc(s, &s->d_lvert, lvert.size())) ||
            (rc = dev_upload(s, s->d_lvert, lvert.data(), lvert.size())) ||
            (rc = dev_upload(s, s->d_gl_start, gl_start.data(), gl_start.size())) ||
            (rc = dev_upload(s, s->d_linc_start, linc_start.data(), linc_start.size())) ||
            (rc = dev_upload(s, s->d_linc, linc.data(), linc.size())) ||
            (rc = dev_upload(s, s->d_lvstart, lvstart.data(), lvstart.size())) ||
            (rc = dev_upload(s, s->d_lvrec, lvrec.data(), lvrec.size())))
            return rc;
    }
    // restraint work: per-group partials, scalars
    s->has_images = false;
    for (const FaceClass &C : s->P.classes)
        for (uint8_t h : C.has_img) s->has_images = s->has_images || h;
    if ((rc = dev_alloc(s, &s->d_rs, 8)) || (rc = dev_alloc(s, &s->d_rpart, 2 * (size_t)s->npartial)) ||
        (rc = dev_alloc(s, &s->d_step, 1)) || (rc = dev_alloc(s, &s->d_fin, 1)) || (rc = dev_alloc(s, &s->d_nan, 1)))
        return rc;
    // face work tables
    std::vector<FaceItem> items;
    for (size_t i = 0; i < s->launch_order.size(); i++) {
        const int c = s->launch_order[i];
        const FaceClass &C = s->P.classes[c];
        for (int g = 0; g < face_groups(C.nf()); g++)
            items.push_back(FaceItem{c, g, s->block_start[i] + g, kind_of(C.ncp, C.npt)});
    }
    if ((rc = dev_alloc(s, &s->d_fa_tab, s->P.classes.size())) || (rc = dev_alloc(s, &s->d_fa_aux, s->P.classes.size())) ||
        (rc = dev_alloc(s, &s->d_items, items.size())) ||
        (rc = dev_upload(s, s->d_items, items.data(), items.size())))
        return rc;
    s->fa_dirty = true;
    const int zero_step = 0;
    const double zeros8[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if ((rc = dev_upload(s, s->d_step, &zero_step, 1)) || (rc = dev_upload(s, s->d_fin, &zero_step, 1)) ||
        (rc = dev_upload(s, s->d_nan, &zero_step, 1)) ||
        (rc = dev_upload(s, s->d_rs, zeros8, 8)))
        return rc;
    if ((rc = sync(s))) return rc;
    s->planned = true;
    return UMCP_OK;
}

int umcp_set_face_c0(umcp_surface *s, const double *c0)
{
    int rc = check_ready(s, false);
    if (rc) return rc;
    if (!c0) return fail(s, UMCP_ERR_ARG, "null c0");
    std::memcpy(s->face_c0.data(), c0, sizeof(double) * s->P.nfaces);
    for (size_t c = 0; c < s->P.classes.size(); c++) {
        const FaceClass &C = s->P.classes[c];
        std::vector<double> v(C.nf());
        for (int f = 0; f < C.nf(); f++) v[f] = c0[C.face_id[f]];
        if ((rc = dev_upload(s, s->cls[c].c0, v.data(), v.size()))) return rc;
    }
    return sync(s);
}

int umcp_setg0(umcp_surface *s, const double *r)
{
    int rc = check_ready(s, false);
    if (rc) return rc;
    if (!r) return fail(s, UMCP_ERR_ARG, "null r");
    if ((rc = upload_r(s, r))) return rc;
    for (size_t c = 0; c < s->P.classes.size(); c++) {
        const FaceClass &C = s->P.classes[c];
        FaceArgs a = face_args(s, c, s->d_r);
        hipError_t e = launch_face_metric(a, kind_of(C.ncp, C.npt), 0, s->cls[c].g0, s->cls[c].rurv0,
                                          s->cls[c].ig0, nullptr, 0, s->stream);
        if (e != hipSuccess) return hip_fail(s, e, "metric kernel launch");
    }
    // the reference records V0 right after setg0 on the same positions
    // (src/hd.C:273-279)
    double h_rs[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (!s->has_images) {
        if ((rc = enqueue_volume(s, s->d_r, s->stream))) return rc;
        hipError_t e = hipMemcpyAsync(h_rs, s->d_rs, sizeof h_rs, hipMemcpyDeviceToHost, s->stream);
        if (e != hipSuccess) return hip_fail(s, e, "download volume");
    }
    std::vector<std::vector<double>> g0s(s->P.classes.size());
    for (size_t c = 0; c < s->P.classes.size(); c++) {
        const FaceClass &C = s->P.classes[c];
        g0s[c].resize((size_t)C.nf() * C.npt);
        hipError_t e = hipMemcpyAsync(g0s[c].data(), s->cls[c].g0, g0s[c].size() * sizeof(double),
                                      hipMemcpyDeviceToHost, s->stream);
        if (e != hipSuccess) return hip_fail(s, e, "download g0");
    }
    if ((rc = sync(s))) return rc;
    s->V_setg0 = h_rs[4];
    // lipid packing per triangle from the reference area of each face, in the
    // reference's face order (regular faces, then irregular): one type,
    // add its area per lipid per leaflet until it exceeds A - 65/2
    // (src/interpLib.C:3113-3196); a composition: the reference's rand()
    // draws (src/interpLib.C:3127-3196, 3286-3358)
    const int nfaces = s->P.nfaces;
    const bool comp = s->lipids.ntypes() > 0;
    const int nt = comp ? s->lipids.ntypes() : 1;
    const std::vector<double> one_c0(1, s->prm.lipid_c0);
    const std::vector<double> &tc0 = comp ? s->lipids.c0 : one_c0;
    std::vector<double> Aface(nfaces, 0.0);
    std::vector<char> irr(nfaces, 0);
    double A0 = 0.0;
    for (size_t c = 0; c < s->P.classes.size(); c++) {
        const FaceClass &C = s->P.classes[c];
        const int nf = C.nf();
        const std::vector<double> &g0 = g0s[c];
        for (int f = 0; f < nf; f++) {
            double A = 0.0;
            for (int p = 0; p < C.npt; p++) A += g0[(size_t)p * nf + f] * C.dudv * C.weight[p];
            Aface[C.face_id[f]] = A;
            irr[C.face_id[f]] = C.valence != 6;
        }
        for (int p = 0; p < C.npt; p++)
            for (int f = 0; f < nf; f++) A0 += C.dudv * g0[(size_t)p * nf + f] * C.weight[p];
    }
    // the packing adds one lipid per round until the face's area is covered:
    // refuse compositions that would put more than 1e7 lipids on a face
    {
        double amax = 0.0, amin_apl = comp ? s->lipids.apl[0] : s->prm.lipid_apl;
        for (int f = 0; f < nfaces; f++) amax = std::max(amax, Aface[f]);
        if (comp)
            for (double a : s->lipids.apl) amin_apl = std::min(amin_apl, a);
        if (!(amin_apl > 0.0) || !(amax / amin_apl <= 1e7))
            return fail(s, UMCP_ERR_ARG, "the lipid packing would place more than 1e7 lipids on a face "
                                         "(area per lipid too small for the faces)");
    }
    s->face_lip_o.assign((size_t)nfaces * nt, 0.0);
    s->face_lip_i.assign((size_t)nfaces * nt, 0.0);
    if (comp) {
        GlibcRand gr(s->lipids.seed);
        for (int f = 0; f < nfaces; f++)
            pack_lipids(s->lipids, gr, Aface[f], irr[f], &s->face_lip_i[(size_t)f * nt], &s->face_lip_o[(size_t)f * nt]);
    } else {
        const double apl = s->prm.lipid_apl;
        for (int f = 0; f < nfaces; f++) {
            double packed = 0.0, L = 0.0;
            while (!(packed > 0 && packed > Aface[f] - 0.5 * 65.0)) {
                L += apl;
                packed += apl;
            }
            s->face_lip_o[f] = s->face_lip_i[f] = L;
        }
    }
    // per face: each leaflet's total, and the lipid energy's (lw, ls, lv)
    for (size_t c = 0; c < s->P.classes.size(); c++) {
        const FaceClass &C = s->P.classes[c];
        const int nf = C.nf();
        std::vector<double> lw(nf), ls(nf), lv(nf);
        for (int f = 0; f < nf; f++) {
            const int gf = C.face_id[f];
            const double *Lo = &s->face_lip_o[(size_t)gf * nt], *Li = &s->face_lip_i[(size_t)gf * nt];
            double so = 0.0, si = 0.0, W = 0.0, Ws = 0.0;
            for (int x = 0; x < nt; x++) {
                so += Lo[x];
                si += Li[x];
                W += Lo[x] + Li[x];
                Ws += Lo[x] * tc0[x] - Li[x] * tc0[x];
            }
            const double sb = W > 0 ? Ws / W : 0.0;
            double V = 0.0;
            for (int x = 0; x < nt; x++) {
                const double d_o = tc0[x] - sb, d_i = -tc0[x] - sb;
                V += Lo[x] * d_o * d_o + Li[x] * d_i * d_i;
            }
            s->face_Lo[gf] = so;
            s->face_Li[gf] = si;
            lw[f] = W;
            ls[f] = sb;
            lv[f] = V;
        }
        if ((rc = dev_upload(s, s->cls[c].lw, lw.data(), nf)) || (rc = dev_upload(s, s->cls[c].ls, ls.data(), nf)) ||
            (rc = dev_upload(s, s->cls[c].lv, lv.data(), nf)))
            return rc;
    }
    s->A0_setg0 = A0;
    if ((rc = sync(s))) return rc;
    s->g0_set = true;
    s->fa_dirty = true;   // restraint targets may default to this state
    return UMCP_OK;
}

static int eval_host(umcp_surface *s, const double *r, double *E, double *g, bool accumulate)
{
    int rc = check_ready(s);
    if (rc) return rc;
    if (!r) return fail(s, UMCP_ERR_ARG, "null r");
    if ((rc = upload_r(s, r))) return rc;
    GatherArgs ga{};
    ga.out_mode = g ? 0 : 2;
    ga.g = s->d_g;
    ga.E_out = s->d_scalar;
    if ((rc = enqueue_eval(s, s->d_r, s->stream, ga))) return rc;
    const size_t n = 3 * (size_t)s->T.nv + 3;
    std::vector<double> gh;
    double Eh = 0.0;
    hipError_t e = hipMemcpyAsync(&Eh, s->d_scalar, sizeof(double), hipMemcpyDeviceToHost, s->stream);
    if (e != hipSuccess) return hip_fail(s, e, "download E");
    if (g) {
        gh.resize(n);
        e = hipMemcpyAsync(gh.data(), s->d_g, n * sizeof(double), hipMemcpyDeviceToHost, s->stream);
        if (e != hipSuccess) return hip_fail(s, e, "download g");
    }
    if ((rc = sync(s))) return rc;
    if ((rc = nan_status(s, false))) return rc;
    if (E) *E = Eh;
    if (g) {
        if (accumulate)
            for (size_t i = 0; i < n; i++) g[i] += gh[i];
        else
            std::memcpy(g, gh.data(), n * sizeof(double));
    }
    return UMCP_OK;
}

int umcp_energy(umcp_surface *s, const double *r, double *E)
{
    if (!E) return fail(s, UMCP_ERR_ARG, "null E");
    return eval_host(s, r, E, nullptr, false);
}

int umcp_grad(umcp_surface *s, const double *r, double *g)
{
    if (!g) return fail(s, UMCP_ERR_ARG, "null g");
    return eval_host(s, r, nullptr, g, true);
}

int umcp_energy_grad(umcp_surface *s, const double *r, double *E, double *g)
{
    if (!g || !E) return fail(s, UMCP_ERR_ARG, "null output");
    return eval_host(s, r, E, g, false);
}

namespace {

// faces in the reference's order as (class, index), the classes' kernel
// kinds, and every vertex's faces the way computeModifiedVertices lists them
// (src/interpLib.C:4040-4088: faces ascending, one entry per control-point
// occurrence); built once per surface
int face_energy_setup(umcp_surface *s)
{
    if (s->fe_ready) return UMCP_OK;
    const Plan &P = s->P;
    std::vector<int2> all((size_t)P.nfaces);
    for (size_t c = 0; c < P.classes.size(); c++)
        for (int i = 0; i < P.classes[c].nf(); i++) all[(size_t)P.classes[c].face_id[i]] = make_int2((int)c, i);
    std::vector<int> kinds;
    for (const FaceClass &C : P.classes) kinds.push_back(kind_of(C.ncp, C.npt));
    const int nv = s->T.nv;
    std::vector<int> cnt((size_t)nv + 1, 0);
    for (int f = 0; f < P.nfaces; f++) {
        const FaceClass &C = P.classes[(size_t)all[(size_t)f].x];
        for (int x = 0; x < C.ncp; x++) cnt[(size_t)C.cp[(size_t)x * C.nf() + all[(size_t)f].y] + 1]++;
    }
    for (int v = 0; v < nv; v++) cnt[(size_t)v + 1] += cnt[(size_t)v];
    s->fe_vstart = cnt;
    s->fe_vfaces.assign((size_t)cnt[(size_t)nv], 0);
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    for (int f = 0; f < P.nfaces; f++) {
        const FaceClass &C = P.classes[(size_t)all[(size_t)f].x];
        for (int x = 0; x < C.ncp; x++) {
            const int v = C.cp[(size_t)x * C.nf() + all[(size_t)f].y];
            s->fe_vfaces[(size_t)fill[(size_t)v]++] = f;
        }
    }
    s->fe_items = all;
    int rc;
    if ((rc = dev_alloc(s, &s->d_fe_kinds, kinds.size())) || (rc = dev_alloc(s, &s->d_fe_items, all.size())) ||
        (rc = dev_alloc(s, &s->d_fe_sub, (size_t)kFeSub)) || (rc = dev_alloc(s, &s->d_fe_out, 4 * all.size())) ||
        (rc = dev_upload(s, s->d_fe_kinds, kinds.data(), kinds.size())) ||
        (rc = dev_upload(s, s->d_fe_items, all.data(), all.size())))
        return rc;
    s->fe_ready = true;
    return UMCP_OK;
}

}  // namespace

}  // extern "C"

// the numerator of every face's curvature-energy density and its area
// (sum dudv g0 w), global face order: penergy's face_energy_density and
// face_area (src/interpLib.C:3612-3613, 3850-3851)
int umcp::face_density(umcp_surface *s, const double *r, double *num, double *area0)
{
    int rc = check_ready(s);
    if (rc) return rc;
    if ((rc = face_energy_setup(s)) || (rc = upload_r(s, r)) || (rc = sync_face_table(s, s->stream))) return rc;
    const int n = s->P.nfaces;
    hipError_t e = launch_face_energies(s->d_fa_tab, s->d_fe_kinds, (int)s->P.classes.size(), s->d_r,
                                        s->d_fe_items, n, s->d_fe_out, s->stream);
    if (e != hipSuccess) return hip_fail(s, e, "face energy launch");
    e = hipMemcpyAsync(num, s->d_fe_out + 3 * (size_t)n, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost,
                       s->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(area0, s->d_fe_out + 2 * (size_t)n, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost,
                           s->stream);
    if (e != hipSuccess) return hip_fail(s, e, "download face densities");
    return sync(s);
}

extern "C" {

int umcp_face_energies(umcp_surface *s, const double *r, double *E, double *A, double *A0)
{
    int rc = check_ready(s);
    if (rc) return rc;
    if (!r || !E) return fail(s, UMCP_ERR_ARG, "null argument");
    if ((rc = face_energy_setup(s)) || (rc = upload_r(s, r)) || (rc = sync_face_table(s, s->stream))) return rc;
    const int n = s->P.nfaces;
    hipError_t e = launch_face_energies(s->d_fa_tab, s->d_fe_kinds, (int)s->P.classes.size(), s->d_r,
                                        s->d_fe_items, n, s->d_fe_out, s->stream);
    if (e != hipSuccess) return hip_fail(s, e, "face energy launch");
    e = hipMemcpyAsync(E, s->d_fe_out, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess && A)
        e = hipMemcpyAsync(A, s->d_fe_out + n, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess && A0)
        e = hipMemcpyAsync(A0, s->d_fe_out + 2 * (size_t)n, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost,
                           s->stream);
    if (e != hipSuccess) return hip_fail(s, e, "download face energies");
    return sync(s);
}

namespace {

// energy, area and reference area of the faces a vertex moves, each summed in
// computeModifiedVertices' order
int vertex_sums(umcp_surface *s, const double *r, int vertex, double out[3])
{
    int rc = check_ready(s);
    if (rc) return rc;
    if (!r) return fail(s, UMCP_ERR_ARG, "null argument");
    if (vertex < 0 || vertex >= s->T.nv) return fail(s, UMCP_ERR_ARG, "vertex index out of range");
    if ((rc = face_energy_setup(s)) || (rc = upload_r(s, r)) || (rc = sync_face_table(s, s->stream))) return rc;
    const int lo = s->fe_vstart[(size_t)vertex], n = s->fe_vstart[(size_t)vertex + 1] - lo;
    if (n > kFeSub) return fail(s, UMCP_ERR_ARG, "vertex is a control point of too many faces");
    std::vector<int2> items((size_t)n);
    for (int i = 0; i < n; i++) items[(size_t)i] = s->fe_items[(size_t)s->fe_vfaces[(size_t)(lo + i)]];
    if ((rc = dev_upload(s, s->d_fe_sub, items.data(), items.size()))) return rc;
    hipError_t e = launch_face_energies(s->d_fa_tab, s->d_fe_kinds, (int)s->P.classes.size(), s->d_r, s->d_fe_sub,
                                        n, s->d_fe_out, s->stream);
    if (e != hipSuccess) return hip_fail(s, e, "face energy launch");
    std::vector<double> h(3 * (size_t)n);
    e = hipMemcpyAsync(h.data(), s->d_fe_out, sizeof(double) * 3 * (size_t)n, hipMemcpyDeviceToHost, s->stream);
    if (e != hipSuccess) return hip_fail(s, e, "download face energies");
    if ((rc = sync(s))) return rc;
    for (int k = 0; k < 3; k++) {
        double sum = 0.0;
        for (int i = 0; i < n; i++) sum += h[(size_t)k * n + i];
        out[k] = sum;
    }
    return UMCP_OK;
}

}  // namespace

int umcp_energy_vertex(umcp_surface *s, const double *r, int vertex, double *E)
{
    // surface::energy with do_vertex >= 0 (src/interpLib.C:4262-4295): no
    // global terms
    if (!E) return fail(s, UMCP_ERR_ARG, "null argument");
    double out[3];
    int rc = vertex_sums(s, r, vertex, out);
    if (rc == UMCP_OK) *E = out[0];
    return rc;
}

int umcp_area_vertex(umcp_surface *s, const double *r, int vertex, double *A, double *A0)
{
    // surface::area with do_vertex >= 0 (src/interpLib.C:4420-4560)
    if (!A || !A0) return fail(s, UMCP_ERR_ARG, "null argument");
    double out[3];
    int rc = vertex_sums(s, r, vertex, out);
    if (rc == UMCP_OK) {
        *A = out[1];
        *A0 = out[2];
    }
    return rc;
}

int umcp_area(umcp_surface *s, const double *r, double *A, double *A0)
{
    int rc = check_ready(s);
    if (rc) return rc;
    if (!r || !A || !A0) return fail(s, UMCP_ERR_ARG, "null argument");
    if ((rc = upload_r(s, r))) return rc;
    int base = 0;
    for (size_t c = 0; c < s->P.classes.size(); c++) {
        const FaceClass &C = s->P.classes[c];
        FaceArgs a = face_args(s, c, s->d_r);
        hipError_t e = launch_face_metric(a, kind_of(C.ncp, C.npt), 1, nullptr, nullptr, nullptr, s->d_partial,
                                          base, s->stream);
        base += metric_blocks(C.nf());
        if (e != hipSuccess) return hip_fail(s, e, "metric kernel launch");
    }
    GatherArgs ga{};
    ga.nv = s->T.nv;
    ga.partial = s->d_partial;
    ga.npartial = base;
    ga.out_mode = 2;
    ga.sums = s->d_scalar;
    hipError_t e = launch_vertex_gather(ga, s->stream);
    if (e != hipSuccess) return hip_fail(s, e, "gather kernel launch");
    double h[4];
    e = hipMemcpyAsync(h, s->d_scalar, 4 * sizeof(double), hipMemcpyDeviceToHost, s->stream);
    if (e != hipSuccess) return hip_fail(s, e, "download area");
    if ((rc = sync(s))) return rc;
    *A = h[0];
    *A0 = h[1];
    return UMCP_OK;
}

int umcp_set_restraints(umcp_surface *s, const umcp_restraints *rs)
{
    if (!s || !rs) return UMCP_ERR_ARG;
    if (!(rs->KA_global >= 0.0) || !(rs->KV >= 0.0)) return fail(s, UMCP_ERR_ARG, "restraint constants must be >= 0");
    if (!s->planned) return fail(s, UMCP_ERR_STATE, "plan has not been generated (call umcp_plan)");
    if (rs->KV != 0.0 && s->has_images)
        return fail(s, UMCP_ERR_ARG, "volume restraint needs a closed mesh (this one has periodic images)");
    s->rsp = *rs;
    s->fa_dirty = true;
    return UMCP_OK;
}

int umcp_volume(umcp_surface *s, const double *r, double *V)
{
    int rc = check_ready(s, false);
    if (rc) return rc;
    if (!r || !V) return fail(s, UMCP_ERR_ARG, "null argument");
    if (s->has_images) return fail(s, UMCP_ERR_ARG, "no enclosed volume: the mesh has periodic images");
    if ((rc = upload_r(s, r))) return rc;
    if ((rc = enqueue_volume(s, s->d_r, s->stream))) return rc;
    double h[8];
    hipError_t e = hipMemcpyAsync(h, s->d_rs, sizeof h, hipMemcpyDeviceToHost, s->stream);
    if (e != hipSuccess) return hip_fail(s, e, "download volume");
    if ((rc = sync(s))) return rc;
    *V = h[4];
    return UMCP_OK;
}

int umcp_set_step(umcp_surface *s, int step)
{
    int rc = check_ready(s, false);
    if (rc) return rc;
    if (step < 0) return fail(s, UMCP_ERR_ARG, "negative step");
    if ((rc = dev_upload(s, s->d_step, &step, 1))) return rc;
    return sync(s);
}

int umcp_get_step(umcp_surface *s, int *step)
{
    int rc = check_ready(s, false);
    if (rc) return rc;
    if (!step) return fail(s, UMCP_ERR_ARG, "null step");
    hipError_t e = hipMemcpyAsync(step, s->d_step, sizeof(int), hipMemcpyDeviceToHost, s->stream);
    if (e != hipSuccess) return hip_fail(s, e, "download step");
    return sync(s);
}

int umcp_bd_steps_rng_dev(umcp_surface *s, double *d_r, int nsteps, const umcp_bd_params *bd,
                          unsigned long long seed, double *d_energies, int energies_cap, void *stream)
{
    int rc = check_ready(s);
    if (rc) return rc;
    if (!d_r || !bd || nsteps < 0 || energies_cap < 0) return fail(s, UMCP_ERR_ARG, "bad Brownian arguments");
    hipStream_t st = stream ? (hipStream_t)stream : s->stream;
    for (int k = 0; k < nsteps; k++) {
        GatherArgs ga{};
        ga.out_mode = 1;
        ga.r = d_r;
        ga.mobility_dt = bd->mobility_dt;
        ga.noise_amp = bd->noise_amp;
        ga.energies = d_energies;
        ga.energies_cap = d_energies ? energies_cap : 0;
        ga.noise = kNoisePhilox;
        ga.seed = seed;
        ga.step_ctr = s->d_step;
        if ((rc = enqueue_eval(s, d_r, st, ga, s->d_step))) return rc;
    }
    return UMCP_OK;
}

namespace {

// point mass (src/hd.C:449-455): two leaflets of POPC (MSCALE 5) over the
// reference area, spread over the lattice points of all faces
double effm_mass_per_point(const umcp_surface *s)
{
    const double kAmuPerKg = 1.0 / 1.661e-27;
    const double mass_per_lipid = 5 * 1000 * 760.09 / 6.022e23 / 1000;
    const double area_per_lipid = 65.35;
    return kAmuPerKg * 2 * s->A0_setg0 / s->P.nfaces / umcp::effm_points_per_face() * mass_per_lipid /
           area_per_lipid;
}

// drop an earlier build (either form)
void effm_release(umcp_surface *s)
{
    (void)hipStreamSynchronize(s->stream);
    for (void *p : s->effm_allocs) {
        auto it = std::find(s->allocations.begin(), s->allocations.end(), p);
        if (it == s->allocations.end()) continue;   // carved from the arena
        (void)hipFree(p);
        s->allocations.erase(it);
    }
    s->effm_allocs.clear();
    s->effm_ready = false;
    s->effm_solve = false;
    s->d_effm = nullptr;
    s->d_effm_rowptr = nullptr;
    s->d_effm_col = nullptr;
    s->d_effm_val = nullptr;
    s->d_ell_col = s->d_ell_perm = nullptr;
    s->d_ell_val = s->d_dinv = s->d_cheb_x = s->d_cheb_r = s->d_cheb_d0 = s->d_cheb_d1 = nullptr;
    s->d_pp_next = s->d_mp = nullptr;
}

// x = Mi b of the scalable form: the Chebyshev solve of M x = b, K
// iterations (b and x device arrays [nv][3] in vertex order): K - 1 full
// steps, then x_K = x_{K-1} + d_{K-1} written back in vertex order
hipError_t effm_solve(umcp_surface *s, const double *b, double *x, hipStream_t st)
{
    using namespace umcp;
    ChebArgs a{};
    a.n = s->T.nv;
    a.width = s->effm_ell_width;
    a.col = s->d_ell_col;
    a.val = s->d_ell_val;
    a.dinv = s->d_dinv;
    a.perm = s->d_ell_perm;
    a.x = s->d_cheb_x;
    a.r = s->d_cheb_r;
    hipError_t e = launch_cheb_init(a, b, s->d_cheb_d0, s->cheb_inv_theta, st);
    double *d[2] = {s->d_cheb_d0, s->d_cheb_d1};
    const size_t K = s->cheb_c1.size();
    for (size_t k = 0; k + 1 < K && e == hipSuccess; k++)
        e = launch_cheb_step(a, d[k & 1], d[(k + 1) & 1], s->cheb_c1[k], s->cheb_c2[k], st);
    if (e == hipSuccess) e = launch_cheb_final(a, d[(K - 1) & 1], x, st);
    return e;
}

// out = Mi v on the device for either form (v, out [nv][3], distinct)
hipError_t effm_apply_dev(umcp_surface *s, const double *v, double *out, hipStream_t st)
{
    using namespace umcp;
    if (s->effm_solve) return effm_solve(s, v, out, st);
    EffmArgs ea{};
    ea.nv = s->T.nv;
    ea.Mi = s->d_effm;
    ea.rowptr = s->d_effm_rowptr;
    ea.col = s->d_effm_col;
    ea.val = s->d_effm_val;
    ea.mp = out;
    return launch_effm_product(ea, v, st);
}

}  // namespace

int umcp_effm_build(umcp_surface *s, long *nnz, double *point_mass)
{
    int rc = check_ready(s);
    if (rc) return rc;
    const int n = s->T.nv;
    if (n > UMCP_EFFM_MAX_NV)
        return fail(s, UMCP_ERR_ARG, "the dense effective-mass inverse holds nv x nv doubles: at most "
                                     "UMCP_EFFM_MAX_NV vertices (umcp_effm_build_sparse has no limit)");
    const double m = effm_mass_per_point(s);
    std::vector<int> rowptr, col;
    std::vector<double> val;
    std::string err;
    if (effm_mass_csr(s->T, s->P, m, rowptr, col, val, err)) return fail(s, UMCP_ERR_BAD_MESH, err);
    // replace an earlier build
    effm_release(s);
    const size_t before = s->allocations.size();
    if ((rc = dev_alloc(s, &s->d_effm, (size_t)n * n)) || (rc = dev_alloc(s, &s->d_pp_next, 3 * (size_t)n)) ||
        (rc = dev_alloc(s, &s->d_mp, 3 * (size_t)n)))
        return rc;
    s->effm_allocs.assign(s->allocations.begin() + (long)before, s->allocations.end());
    // set-up scratch, freed below
    int *d_rowptr = nullptr, *d_col = nullptr, *d_info = nullptr, *d_cnt = nullptr;
    double *d_val = nullptr;
    hipError_t e = hipMalloc(&d_rowptr, sizeof(int) * rowptr.size());
    if (e == hipSuccess) e = hipMalloc(&d_col, sizeof(int) * col.size());
    if (e == hipSuccess) e = hipMalloc(&d_val, sizeof(double) * val.size());
    if (e == hipSuccess) e = hipMalloc(&d_info, sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&d_cnt, sizeof(int) * (size_t)n);
    auto cleanup = [&]() {
        (void)hipStreamSynchronize(s->stream);
        (void)hipFree(d_rowptr);
        (void)hipFree(d_col);
        (void)hipFree(d_val);
        (void)hipFree(d_info);
        (void)hipFree(d_cnt);
    };
    if (e != hipSuccess) {
        cleanup();
        return fail(s, UMCP_ERR_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    e = hipMemcpyAsync(d_rowptr, rowptr.data(), sizeof(int) * rowptr.size(), hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_col, col.data(), sizeof(int) * col.size(), hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_val, val.data(), sizeof(double) * val.size(), hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) e = hipMemsetAsync(s->d_effm, 0, sizeof(double) * (size_t)n * n, s->stream);
    if (e == hipSuccess) e = launch_effm_scatter(n, d_rowptr, d_col, d_val, s->d_effm, s->stream);
    if (e != hipSuccess) {
        cleanup();
        return hip_fail(s, e, "effective mass upload");
    }
    // inverse by Cholesky on the device: rocSOLVER's potrf + potri, the
    // LAPACK pair the reference calls (src/modes.C:1827-1847)
    rocblas_handle h = nullptr;
    int info_f = -1, info_i = -1;
    rocblas_status bs = rocblas_create_handle(&h);
    if (bs == rocblas_status_success) bs = rocblas_set_stream(h, s->stream);
    if (bs == rocblas_status_success) bs = rocsolver_dpotrf(h, rocblas_fill_upper, n, s->d_effm, n, d_info);
    if (bs == rocblas_status_success)
        e = hipMemcpyAsync(&info_f, d_info, sizeof(int), hipMemcpyDeviceToHost, s->stream);
    if (bs == rocblas_status_success && e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (bs == rocblas_status_success && e == hipSuccess && info_f == 0)
        bs = rocsolver_dpotri(h, rocblas_fill_upper, n, s->d_effm, n, d_info);
    if (bs == rocblas_status_success && e == hipSuccess && info_f == 0) {
        e = hipMemcpyAsync(&info_i, d_info, sizeof(int), hipMemcpyDeviceToHost, s->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    }
    if (h) (void)rocblas_destroy_handle(h);
    if (bs != rocblas_status_success || e != hipSuccess) {
        cleanup();
        return fail(s, UMCP_ERR_DEVICE, std::string("effective mass inverse: ") +
                                            (e != hipSuccess ? hipGetErrorString(e) : "rocSOLVER call failed"));
    }
    if (info_f != 0 || info_i != 0) {
        cleanup();
        return fail(s, UMCP_ERR_BAD_MESH, "effective mass matrix is not positive definite (pivot " +
                                              std::to_string(info_f ? info_f : info_i) + ")");
    }
    std::vector<int> cnt((size_t)n);
    e = launch_effm_symmetrize(n, s->d_effm, s->stream);
    if (e == hipSuccess) e = launch_effm_truncate(n, s->d_effm, 1e-20, d_cnt, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(cnt.data(), d_cnt, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost, s->stream);
    cleanup();
    if (e != hipSuccess) return hip_fail(s, e, "effective mass truncation");
    long kept = 0;
    for (int c : cnt) kept += c;
    s->d_effm_rowptr = nullptr;
    s->d_effm_col = nullptr;
    s->d_effm_val = nullptr;
    // CSR (10 bytes per kept entry: 16-bit columns) when that reads less than
    // the dense array (8 bytes per entry): the inverse of a large mesh's mass
    // matrix decays below the truncation far from the diagonal
    if (5 * kept < 4 * (long)n * n && kept < (long)INT32_MAX) {
        std::vector<int> rp((size_t)n + 1, 0);
        for (int i = 0; i < n; i++) rp[(size_t)i + 1] = rp[(size_t)i] + cnt[(size_t)i];
        const size_t before2 = s->allocations.size();
        if ((rc = dev_alloc(s, &s->d_effm_rowptr, rp.size())) || (rc = dev_alloc(s, &s->d_effm_col, (size_t)kept)) ||
            (rc = dev_alloc(s, &s->d_effm_val, (size_t)kept)))
            return rc;
        s->effm_allocs.insert(s->effm_allocs.end(), s->allocations.begin() + (long)before2, s->allocations.end());
        e = hipMemcpyAsync(s->d_effm_rowptr, rp.data(), sizeof(int) * rp.size(), hipMemcpyHostToDevice, s->stream);
        if (e == hipSuccess)
            e = launch_effm_compact(n, s->d_effm, s->d_effm_rowptr, s->d_effm_col, s->d_effm_val, s->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        if (e != hipSuccess) return hip_fail(s, e, "effective mass compaction");
        // the dense array is no longer needed
        (void)hipFree(s->d_effm);
        s->allocations.erase(std::find(s->allocations.begin(), s->allocations.end(), (void *)s->d_effm));
        s->effm_allocs.erase(std::find(s->effm_allocs.begin(), s->effm_allocs.end(), (void *)s->d_effm));
        s->device_bytes -= sizeof(double) * (size_t)n * n;
        s->d_effm = nullptr;
    }
    s->effm_nnz = kept;
    s->effm_point_mass = m;
    s->effm_ready = true;
    if (nnz) *nnz = kept;
    if (point_mass) *point_mass = m;
    return UMCP_OK;
}

int umcp_effm_build_sparse(umcp_surface *s, double tol, long *nnz, double *point_mass, int *iterations)
{
    int rc = check_ready(s);
    if (rc) return rc;
    if (tol == 0.0) tol = 1e-12;
    if (!(tol >= 1e-14 && tol <= 1e-3)) return fail(s, UMCP_ERR_ARG, "solve tolerance outside [1e-14, 1e-3]");
    const int n = s->T.nv;
    const double m = effm_mass_per_point(s);
    std::vector<int> rowptr, col;
    std::vector<double> val;
    std::string err;
    if (effm_mass_csr(s->T, s->P, m, rowptr, col, val, err)) return fail(s, UMCP_ERR_BAD_MESH, err);
    double lo = 0.0, hi = 0.0;
    if (effm_jacobi_interval(rowptr, col, val, lo, hi))
        return fail(s, UMCP_ERR_BAD_MESH, "effective mass matrix is not positive definite");
    int width = 0;
    std::vector<int> ell_col;
    std::vector<double> ell_val, dinv;
    const std::vector<int> perm = effm_rcm_order(rowptr, col);
    effm_ell(rowptr, col, val, perm, width, ell_col, ell_val, dinv);
    effm_release(s);
    const size_t before = s->allocations.size();
    const size_t n3 = 3 * (size_t)n;
    if ((rc = dev_alloc(s, &s->d_pp_next, n3)) || (rc = dev_alloc(s, &s->d_mp, n3)) ||
        (rc = dev_alloc(s, &s->d_ell_col, ell_col.size())) || (rc = dev_alloc(s, &s->d_ell_val, ell_val.size())) ||
        (rc = dev_alloc(s, &s->d_dinv, (size_t)n)) || (rc = dev_alloc(s, &s->d_ell_perm, (size_t)n)) ||
        (rc = dev_alloc(s, &s->d_cheb_x, n3)) || (rc = dev_alloc(s, &s->d_cheb_r, n3)) ||
        (rc = dev_alloc(s, &s->d_cheb_d0, n3)) || (rc = dev_alloc(s, &s->d_cheb_d1, n3))) {
        s->effm_allocs.assign(s->allocations.begin() + (long)before, s->allocations.end());
        effm_release(s);
        return rc;
    }
    s->effm_allocs.assign(s->allocations.begin() + (long)before, s->allocations.end());
    hipError_t e = hipMemcpyAsync(s->d_ell_col, ell_col.data(), sizeof(int) * ell_col.size(), hipMemcpyHostToDevice,
                                  s->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(s->d_ell_val, ell_val.data(), sizeof(double) * ell_val.size(), hipMemcpyHostToDevice,
                           s->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(s->d_dinv, dinv.data(), sizeof(double) * dinv.size(), hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(s->d_ell_perm, perm.data(), sizeof(int) * perm.size(), hipMemcpyHostToDevice, s->stream);
    s->effm_ell_width = width;
    s->effm_solve = true;
    s->cheb_lo = lo;
    s->cheb_hi = hi;
    // the iteration count of the Chebyshev bound, checked on a pseudo-random
    // right-hand side: the residual of the device solve, formed on the host
    // from the CSR, must fall below tol (else 1.5 times the iterations, up to
    // four tries)
    std::vector<double> b(n3), x(n3);
    unsigned long long z = 0x2545F4914F6CDD1Dull;
    double bn = 0.0;
    for (double &q : b) {
        z ^= z << 13, z ^= z >> 7, z ^= z << 17;
        q = (double)(z >> 11) * (1.0 / 9007199254740992.0) - 0.5;
        bn += q * q;
    }
    bn = std::sqrt(bn);
    int iters = effm_cheb_iterations(lo, hi, tol);
    double res = 0.0;
    for (int attempt = 0; attempt < 4 && e == hipSuccess; attempt++) {
        effm_cheb_coeffs(lo, hi, iters, s->cheb_c1, s->cheb_c2, s->cheb_inv_theta);
        e = hipMemcpyAsync(s->d_pp_next, b.data(), sizeof(double) * n3, hipMemcpyHostToDevice, s->stream);
        if (e == hipSuccess) e = effm_solve(s, s->d_pp_next, s->d_mp, s->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(x.data(), s->d_mp, sizeof(double) * n3, hipMemcpyDeviceToHost, s->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        if (e != hipSuccess) break;
        double rn = 0.0;
        for (int i = 0; i < n; i++)
            for (int c = 0; c < 3; c++) {
                double q = b[3 * (size_t)i + c];
                for (int k = rowptr[i]; k < rowptr[i + 1]; k++) q -= val[k] * x[3 * (size_t)col[k] + c];
                rn += q * q;
            }
        res = std::sqrt(rn) / bn;
        if (res <= tol) break;
        iters = (iters * 3 + 1) / 2;
    }
    if (e != hipSuccess) {
        effm_release(s);
        return hip_fail(s, e, "effective mass solve set-up");
    }
    if (!(res <= tol)) {
        effm_release(s);
        return fail(s, UMCP_ERR_BAD_MESH, "Chebyshev solve of the effective mass did not reach the tolerance "
                                          "(residual " + std::to_string(res) + ")");
    }
    s->effm_nnz = (long)val.size();
    s->effm_point_mass = m;
    s->effm_ready = true;
    if (nnz) *nnz = s->effm_nnz;
    if (point_mass) *point_mass = m;
    if (iterations) *iterations = (int)s->cheb_c1.size();
    return UMCP_OK;
}

int umcp_effm_apply(umcp_surface *s, const double *v, double *out)
{
    if (!s || !v || !out) return s ? fail(s, UMCP_ERR_ARG, "null argument") : UMCP_ERR_ARG;
    if (!s->effm_ready) return fail(s, UMCP_ERR_STATE, "effective mass not built (call umcp_effm_build)");
    if (hipSetDevice(s->device) != hipSuccess) return fail(s, UMCP_ERR_DEVICE, "hipSetDevice failed");
    const size_t n3 = 3 * (size_t)s->T.nv;
    double *d_v = nullptr, *d_o = nullptr;
    hipError_t e = hipMalloc(&d_v, sizeof(double) * n3);
    if (e == hipSuccess) e = hipMalloc(&d_o, sizeof(double) * n3);
    if (e == hipSuccess) e = hipMemcpyAsync(d_v, v, sizeof(double) * n3, hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) e = effm_apply_dev(s, d_v, d_o, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out, d_o, sizeof(double) * n3, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    (void)hipFree(d_v);
    (void)hipFree(d_o);
    if (e != hipSuccess) return hip_fail(s, e, "effective mass product");
    return UMCP_OK;
}

int umcp_effm_get(umcp_surface *s, double *Minv)
{
    if (!s || !Minv) return UMCP_ERR_ARG;
    if (!s->effm_ready) return fail(s, UMCP_ERR_STATE, "effective mass not built (call umcp_effm_build)");
    if (s->effm_solve)
        return fail(s, UMCP_ERR_STATE, "the sparse form keeps no inverse (products: umcp_effm_apply)");
    const size_t n = (size_t)s->T.nv;
    if (s->d_effm) {
        hipError_t e = hipMemcpy(Minv, s->d_effm, sizeof(double) * n * n, hipMemcpyDeviceToHost);
        if (e != hipSuccess) return hip_fail(s, e, "download effective mass");
        return UMCP_OK;
    }
    std::vector<int> rp(n + 1);
    std::vector<uint16_t> col((size_t)s->effm_nnz);
    std::vector<double> val((size_t)s->effm_nnz);
    hipError_t e = hipMemcpy(rp.data(), s->d_effm_rowptr, sizeof(int) * rp.size(), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(col.data(), s->d_effm_col, sizeof(uint16_t) * col.size(), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(val.data(), s->d_effm_val, sizeof(double) * val.size(), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(s, e, "download effective mass");
    std::memset(Minv, 0, sizeof(double) * n * n);
    for (size_t i = 0; i < n; i++)
        for (int k = rp[i]; k < rp[i + 1]; k++) Minv[i * n + (size_t)col[(size_t)k]] = val[(size_t)k];
    return UMCP_OK;
}

int umcp_effm_steps_dev(umcp_surface *s, double *d_r, double *d_pp, const double *d_xi, int nsteps,
                        const umcp_effm_params *p, unsigned long long seed, double *d_energies, void *stream)
{
    int rc = check_ready(s);
    if (rc) return rc;
    if (!s->effm_ready) return fail(s, UMCP_ERR_STATE, "effective mass not built (call umcp_effm_build)");
    if (!d_r || !d_pp || !p || nsteps < 0) return fail(s, UMCP_ERR_ARG, "null argument");
    if (p->mode != UMCP_EFFM_BD && p->mode != UMCP_EFFM_LD) return fail(s, UMCP_ERR_ARG, "unknown integrator");
    if (!(p->gamma > 0.0) || !(p->dt > 0.0) || !(p->kT >= 0.0))
        return fail(s, UMCP_ERR_ARG, "effective-mass integrator needs gamma > 0, dt > 0, kT >= 0");
    hipStream_t st = stream ? (hipStream_t)stream : s->stream;
    const size_t n3 = 3 * (size_t)s->T.nv;
    const bool lang = p->mode == UMCP_EFFM_LD;
    EffmArgs ea{};
    ea.nv = s->T.nv;
    ea.Mi = s->d_effm;
    ea.rowptr = s->d_effm_rowptr;
    ea.col = s->d_effm_col;
    ea.val = s->d_effm_val;
    ea.pp_next = s->d_pp_next;
    ea.pp_keep = d_pp;
    ea.mp = s->d_mp;
    ea.r = d_r;
    ea.mode = lang ? kEffmLangevin : kEffmBrownian;
    ea.akma = 1.0 / (0.04888 * (1e-12));   // AKMA_TIME, src/units.h
    ea.dt = p->dt;
    hipError_t e = hipSuccess;
    if (lang && nsteps > 0) {
        // the first step's friction product from the caller's momenta; later
        // steps reuse the product the position kernel (or the solve) forms
        e = effm_apply_dev(s, d_pp, s->d_mp, st);
        if (e != hipSuccess) return hip_fail(s, e, "effective-mass product launch");
    }
    for (int k = 0; k < nsteps; k++) {
        GatherArgs ga{};
        ga.out_mode = 3;
        ga.r = d_r;
        ga.energies = d_energies;
        ga.step = k;
        if (d_xi) {
            ga.noise = kNoiseArray;
            ga.xi = d_xi + (size_t)k * n3;
        } else {
            ga.noise = kNoisePhilox;
            ga.seed = seed;
            ga.step_ctr = s->d_step;
        }
        ga.pp_in = d_pp;
        ga.pp_out = s->d_pp_next;
        ga.mp = s->d_mp;
        ga.effm_mode = ea.mode;
        ga.effm_akma = ea.akma;
        ga.effm_dt = p->dt;
        ga.effm_kT = p->kT;
        ga.effm_gamma = p->gamma;
        if ((rc = enqueue_eval(s, d_r, st, ga, s->d_step))) return rc;
        if (s->effm_solve) {
            // mp = Mi pp_next by the solve (the gather above has read the
            // previous mp), then the positions from it
            e = effm_solve(s, s->d_pp_next, s->d_mp, st);
            if (e == hipSuccess) e = launch_effm_position_solved(ea, st);
        } else {
            e = launch_effm_position(ea, st);
        }
        if (e != hipSuccess) return hip_fail(s, e, "effective-mass step launch");
    }
    return UMCP_OK;
}

int umcp_effm_steps(umcp_surface *s, double *r, double *pp, const double *xi, int nsteps, const umcp_effm_params *p,
                    double *energies)
{
    int rc = check_ready(s);
    if (rc) return rc;
    if (!r || !pp || !xi || nsteps < 0) return fail(s, UMCP_ERR_ARG, "null argument");
    const size_t n3 = 3 * (size_t)s->T.nv;
    double *d_pp = nullptr, *d_xi = nullptr, *d_en = nullptr;
    hipError_t e = hipMalloc(&d_pp, sizeof(double) * n3);
    if (e == hipSuccess && nsteps > 0) e = hipMalloc(&d_xi, sizeof(double) * n3 * (size_t)nsteps);
    if (e == hipSuccess && energies && nsteps > 0) e = hipMalloc(&d_en, sizeof(double) * (size_t)nsteps);
    if (e == hipSuccess) e = hipMemcpyAsync(d_pp, pp, sizeof(double) * n3, hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess && nsteps > 0)
        e = hipMemcpyAsync(d_xi, xi, sizeof(double) * n3 * (size_t)nsteps, hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) rc = upload_r(s, r);
    if (e == hipSuccess && rc == UMCP_OK) rc = umcp_effm_steps_dev(s, s->d_r, d_pp, d_xi, nsteps, p, 0, d_en, nullptr);
    if (e == hipSuccess && rc == UMCP_OK) {
        e = hipMemcpyAsync(r, s->d_r, sizeof(double) * (n3 + 3), hipMemcpyDeviceToHost, s->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(pp, d_pp, sizeof(double) * n3, hipMemcpyDeviceToHost, s->stream);
        if (e == hipSuccess && d_en)
            e = hipMemcpyAsync(energies, d_en, sizeof(double) * (size_t)nsteps, hipMemcpyDeviceToHost, s->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    }
    (void)hipFree(d_pp);
    (void)hipFree(d_xi);
    (void)hipFree(d_en);
    if (e != hipSuccess) return hip_fail(s, e, "effective-mass steps");
    return rc;
}

int umcp_noise(unsigned long long seed, int step, int n, double *d_out, void *stream)
{
    if (!d_out || n < 0 || step < 0) return UMCP_ERR_ARG;
    if (hipSetDevice(g_device) != hipSuccess) return UMCP_ERR_DEVICE;
    if (launch_noise(seed, step, n, d_out, (hipStream_t)stream) != hipSuccess) return UMCP_ERR_DEVICE;
    return UMCP_OK;
}

int umcp_energy_grad_dev(umcp_surface *s, const double *d_r, double *d_E, double *d_g, void *stream)
{
    int rc = check_ready(s);
    if (rc) return rc;
    GatherArgs ga{};
    ga.out_mode = d_g ? 0 : 2;
    ga.g = d_g;
    ga.E_out = d_E;
    return enqueue_eval(s, d_r, stream ? (hipStream_t)stream : s->stream, ga);
}

int umcp_bd_steps_dev(umcp_surface *s, double *d_r, const double *d_xi, int nsteps, const umcp_bd_params *bd,
                      double *d_energies, void *stream)
{
    int rc = check_ready(s);
    if (rc) return rc;
    if (!bd || nsteps < 0) return fail(s, UMCP_ERR_ARG, "bad Brownian parameters");
    hipStream_t st = stream ? (hipStream_t)stream : s->stream;
    const size_t n3 = 3 * (size_t)s->T.nv;
    for (int k = 0; k < nsteps; k++) {
        GatherArgs ga{};
        ga.out_mode = 1;
        ga.r = d_r;
        ga.xi = d_xi + (size_t)k * n3;
        ga.mobility_dt = bd->mobility_dt;
        ga.noise_amp = bd->noise_amp;
        ga.energies = d_energies;
        ga.step = k;
        if ((rc = enqueue_eval(s, d_r, st, ga))) return rc;
    }
    return UMCP_OK;
}

int umcp_bd_steps(umcp_surface *s, double *r, const double *xi, int nsteps, const umcp_bd_params *bd,
                  double *energies)
{
    int rc = check_ready(s);
    if (rc) return rc;
    if (!r || !xi || !bd || nsteps < 0) return fail(s, UMCP_ERR_ARG, "null argument");
    const size_t n3 = 3 * (size_t)s->T.nv;
    const int chunk = nsteps < 64 ? (nsteps > 0 ? nsteps : 1) : 64;
    if (s->xi_chunk < chunk) {
        if ((rc = dev_alloc(s, &s->d_xi, n3 * chunk))) return rc;
        s->xi_chunk = chunk;
    }
    if (s->energies_cap < chunk) {
        if ((rc = dev_alloc(s, &s->d_energies, chunk))) return rc;
        s->energies_cap = chunk;
    }
    if ((rc = upload_r(s, r))) return rc;
    for (int k0 = 0; k0 < nsteps; k0 += chunk) {
        const int nk = (nsteps - k0 < chunk) ? nsteps - k0 : chunk;
        if ((rc = dev_upload(s, s->d_xi, xi + (size_t)k0 * n3, n3 * nk))) return rc;
        if ((rc = umcp_bd_steps_dev(s, s->d_r, s->d_xi, nk, bd, s->d_energies, s->stream))) return rc;
        if (energies) {
            hipError_t e = hipMemcpyAsync(energies + k0, s->d_energies, nk * sizeof(double),
                                          hipMemcpyDeviceToHost, s->stream);
            if (e != hipSuccess) return hip_fail(s, e, "download energies");
        }
        if ((rc = sync(s))) return rc;  // the staging buffer is reused by the next chunk
        if ((rc = nan_status(s, false))) return rc;
    }
    hipError_t e = hipMemcpyAsync(r, s->d_r, n3 * sizeof(double), hipMemcpyDeviceToHost, s->stream);
    if (e != hipSuccess) return hip_fail(s, e, "download r");
    return sync(s);
}

int umcp_nan_check(umcp_surface *s)
{
    int rc = check_ready(s, false);
    if (rc) return rc;
    return nan_status(s, true);
}

int umcp_set_launch(umcp_surface *s, int point_split, int flags)
{
    if (!s || !s->planned) return UMCP_ERR_STATE;
    if (const char *why = launch_shape_error(point_split, flags)) return fail(s, UMCP_ERR_ARG, why);
    s->point_split = point_split;
    s->launch_flags = flags;
    return UMCP_OK;
}

int umcp_get_launch(const umcp_surface *s, int *point_split, int *flags)
{
    if (!s || !s->planned) return UMCP_ERR_STATE;
    if (point_split) *point_split = s->point_split;
    if (flags) *flags = s->launch_flags;
    return UMCP_OK;
}

int umcp_profile_begin(umcp_surface *s)
{
    if (!s || !s->planned) return UMCP_ERR_STATE;
    s->prof_on = true;
    s->prof_used = 0;
    s->prof_tick = 0;
    return UMCP_OK;
}

int umcp_profile_sample(umcp_surface *s, int every)
{
    if (!s) return UMCP_ERR_ARG;
    if (every < 1) return fail(s, UMCP_ERR_ARG, "sampling stride must be >= 1");
    s->prof_every = every;
    return UMCP_OK;
}

int umcp_profile_end(umcp_surface *s, int max_cls, double *ms, int *launches)
{
    if (!s || !s->planned) return UMCP_ERR_STATE;
    s->prof_on = false;
    for (int c = 0; c < max_cls; c++) {
        if (ms) ms[c] = 0.0;
        if (launches) launches[c] = 0;
    }
    for (size_t i = 0; i < s->prof_used; i++) {
        hipError_t e = hipEventSynchronize(s->prof_events[2 * i + 1]);
        if (e != hipSuccess) return hip_fail(s, e, "hipEventSynchronize");
        float t = 0.0f;
        e = hipEventElapsedTime(&t, s->prof_events[2 * i], s->prof_events[2 * i + 1]);
        if (e != hipSuccess) return hip_fail(s, e, "hipEventElapsedTime");
        int c = s->prof_class[i];
        if (c < max_cls) {
            if (ms) ms[c] += t;
            if (launches) launches[c] += 1;
        }
    }
    s->prof_used = 0;
    return UMCP_OK;
}

int umcp_class_count(const umcp_surface *s) { return (s && s->prepared) ? (int)s->P.classes.size() : 0; }

int umcp_class_info(const umcp_surface *s, int cls, int *valence, int *ncp, int *npt, int *nf)
{
    if (!s || !s->prepared || cls < 0 || cls >= (int)s->P.classes.size()) return UMCP_ERR_ARG;
    const FaceClass &C = s->P.classes[cls];
    if (valence) *valence = C.valence;
    if (ncp) *ncp = C.ncp;
    if (npt) *npt = C.npt;
    if (nf) *nf = C.nf();
    return UMCP_OK;
}

int umcp_class_table(const umcp_surface *s, int cls, double *table, double *weights)
{
    if (!s || !s->prepared || cls < 0 || cls >= (int)s->P.classes.size()) return UMCP_ERR_ARG;
    const FaceClass &C = s->P.classes[cls];
    if (table) std::memcpy(table, C.table.data(), C.table.size() * sizeof(double));
    if (weights) std::memcpy(weights, C.weight.data(), C.weight.size() * sizeof(double));
    return UMCP_OK;
}

int umcp_face_points(const umcp_surface *s, int face, int *ncp, int *cp)
{
    if (!s || !s->prepared || face < 0 || face >= s->P.nfaces) return UMCP_ERR_ARG;
    for (const FaceClass &C : s->P.classes)
        for (int f = 0; f < C.nf(); f++)
            if (C.face_id[f] == face) {
                if (ncp) *ncp = C.ncp;
                if (cp)
                    for (int q = 0; q < C.ncp; q++) cp[q] = C.cp[(size_t)q * C.nf() + f];
                return UMCP_OK;
            }
    return UMCP_ERR_ARG;
}

int umcp_face_lipids(const umcp_surface *s, int face, double *Lo, double *Li)
{
    if (!s || !s->g0_set || face < 0 || face >= s->P.nfaces) return UMCP_ERR_ARG;
    if (Lo) *Lo = s->face_Lo[face];
    if (Li) *Li = s->face_Li[face];
    return UMCP_OK;
}

int umcp_lipid_type_count(const umcp_surface *s)
{
    if (!s) return UMCP_ERR_ARG;
    return s->lipids.ntypes() > 0 ? s->lipids.ntypes() : 1;
}

int umcp_face_lipid_types(const umcp_surface *s, int face, double *Lo, double *Li)
{
    if (!s || !s->g0_set || face < 0 || face >= s->P.nfaces || !Lo || !Li) return UMCP_ERR_ARG;
    const size_t nt = s->face_lip_o.size() / (size_t)s->P.nfaces;
    for (size_t x = 0; x < nt; x++) {
        Lo[x] = s->face_lip_o[(size_t)face * nt + x];
        Li[x] = s->face_lip_i[(size_t)face * nt + x];
    }
    return UMCP_OK;
}

int umcp_set_lipids(umcp_surface *s, int ntypes, const double *c0, const double *apl, const double *inner_parts,
                    const double *outer_parts, unsigned int seed)
{
    if (!s) return UMCP_ERR_ARG;
    if (ntypes < 1 || ntypes > kMaxLipidTypes || !c0 || !apl || !inner_parts || !outer_parts)
        return fail(s, UMCP_ERR_ARG, "a composition has 1 to 64 lipid types");
    LipidComposition L;
    double si = 0.0, so = 0.0;
    for (int x = 0; x < ntypes; x++) {
        if (!(apl[x] > 0.0) || !(inner_parts[x] >= 0.0) || !(outer_parts[x] >= 0.0) || !std::isfinite(c0[x]))
            return fail(s, UMCP_ERR_ARG, "lipid types need an area per lipid above 0 and parts of at least 0");
        si += inner_parts[x];
        so += outer_parts[x];
        L.names.push_back("type" + std::to_string(x));
        L.c0.push_back(c0[x]);
        L.apl.push_back(apl[x]);
        L.inner.push_back(inner_parts[x]);
        L.outer.push_back(outer_parts[x]);
    }
    if (!(si > 0.0) || !(so > 0.0)) return fail(s, UMCP_ERR_ARG, "each leaflet needs parts above 0");
    L.seed = seed;
    s->lipids = L;
    return UMCP_OK;
}

int umcp_read_lipid_composition(umcp_surface *s, const char *path, unsigned int seed)
{
    if (!s || !path) return s ? fail(s, UMCP_ERR_ARG, "null path") : UMCP_ERR_ARG;
    LipidComposition L;
    std::string err;
    if (read_lipid_composition(path, seed, L, err)) return fail(s, UMCP_ERR_ARG, err);
    s->lipids = L;
    return UMCP_OK;
}

}  // extern "C"
