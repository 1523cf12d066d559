// Batched evaluation of many independent surfaces (umcp_ensemble_* in
// umcp.h).  The reference runs one surface per process (src/hd.C); replicas
// and parameter sweeps are separate jobs.  Here every member stays resident in
// HBM with its own tables and work buffers, and one step of all members is
// three launches over device-resident work tables:
//
//   restraint pre-pass       area / volume partials and multipliers of the
//                            members that have restraints (same items)
//   faces_batch_kernel       one item per 64-face group of every member
//   vertex_gather_batch      one item per 256-vertex block of every member
//
// An item runs exactly the code the member's own launch runs and writes to
// the same partial slot, so a member's trajectory is the same alone and in
// any mix.  The noise is the member's counter-based stream (seed, step), so
// it does not depend on the mix either.  The step sequence can be captured
// into a hipGraph and replayed; tables and graph are rebuilt when membership
// changes.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/umcp.h"
#include "umcp_kernels.hpp"
#include "umcp_state.hpp"

using namespace umcp;

struct umcp_ensemble {
    struct Member {
        umcp_surface *s;
        double *d_r;
        umcp_bd_params bd;
        unsigned long long seed;
        double *d_energies;
        int energies_cap;
    };
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::vector<Member> members;
    bool dirty = true;
    // launch shape
    int point_split = 4, flags = kLaunchOcc2, face_grid = 0, gather_grid = 0, order = 0;
    // built tables; face items are grouped by energy form (the kernel a member
    // runs alone), one batched launch per non-empty group
    bool mixed = false;
    int feat_start[4] = {0, 0, 0, 0};
    FaceArgs *d_fa = nullptr;
    FaceItem *d_items = nullptr;
    GatherArgs *d_ga = nullptr;
    GatherItem *d_gitems = nullptr;
    int *d_fin = nullptr;         // index in d_items of each restrained member's slot-0 item
    int nitems = 0, ngitems = 0, nrestrained = 0, nfin = 0;
    std::vector<void *> tables;   // allocations of the current build
    // captured step sequence
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_steps = 0;
    hipStream_t graph_stream = nullptr;
    hipStream_t last_stream = nullptr;   // stream of the latest steps (may be the caller's)
};

namespace {

int efail(umcp_ensemble *e, int code, const std::string &msg)
{
    if (e) e->err = msg;
    return code;
}

int ehip(umcp_ensemble *e, hipError_t err, const char *where)
{
    return efail(e, UMCP_ERR_DEVICE, std::string(where) + ": " + hipGetErrorString(err));
}

// waits for every launch that may still read the work tables or a member
hipError_t quiesce(umcp_ensemble *e)
{
    hipError_t err = hipStreamSynchronize(e->stream);
    if (err == hipSuccess && e->last_stream && e->last_stream != e->stream)
        err = hipStreamSynchronize(e->last_stream);
    return err;
}

void drop_graph(umcp_ensemble *e)
{
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->exec = nullptr;
    e->graph = nullptr;
    e->graph_steps = 0;
    e->graph_stream = nullptr;
}

void drop_tables(umcp_ensemble *e)
{
    for (void *p : e->tables) (void)hipFree(p);
    e->tables.clear();
    e->d_fa = nullptr;
    e->d_items = nullptr;
    e->d_ga = nullptr;
    e->d_gitems = nullptr;
    e->d_fin = nullptr;
    e->nitems = e->ngitems = e->nrestrained = e->nfin = 0;
}

template <typename T>
int upload_table(umcp_ensemble *e, T **dst, const std::vector<T> &src)
{
    *dst = nullptr;
    if (src.empty()) return UMCP_OK;
    void *p = nullptr;
    hipError_t err = hipMalloc(&p, src.size() * sizeof(T));
    if (err != hipSuccess) return efail(e, UMCP_ERR_OOM, std::string("hipMalloc: ") + hipGetErrorString(err));
    e->tables.push_back(p);
    err = hipMemcpyAsync(p, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, e->stream);
    if (err != hipSuccess) return ehip(e, err, "upload work table");
    *dst = (T *)p;
    return UMCP_OK;
}

// work tables of the current membership
int build(umcp_ensemble *e)
{
    drop_graph(e);
    hipError_t err = quiesce(e);   // old tables may still be read
    if (err != hipSuccess) return ehip(e, err, "hipStreamSynchronize");
    drop_tables(e);
    e->dirty = false;
    if (e->members.empty()) return UMCP_OK;
    const int prec = e->members[0].s->prm.precision;
    std::vector<int> feat_of(e->members.size());
    for (size_t mi = 0; mi < e->members.size(); mi++) {
        const auto &m = e->members[mi];
        if (m.s->prm.precision != prec)
            return efail(e, UMCP_ERR_ARG, "all members of an ensemble must use the same precision");
        feat_of[mi] = features_of(face_args(m.s, 0, m.d_r).ph);
    }
    e->mixed = prec == UMCP_PREC_MIXED;

    std::vector<FaceArgs> fa;
    std::vector<std::vector<FaceItem>> per_member(e->members.size());
    std::vector<GatherArgs> ga;
    std::vector<GatherItem> gitems;
    int nrestrained = 0;
    for (size_t mi = 0; mi < e->members.size(); mi++) {
        const auto &m = e->members[mi];
        umcp_surface *s = m.s;
        const bool rson = restraints_on(s);
        if (rson) nrestrained++;
        std::vector<int> fa_of_class(s->P.classes.size());
        for (size_t c = 0; c < s->P.classes.size(); c++) {
            FaceArgs a = face_args(s, c, m.d_r);
            a.step_ctr = s->d_step;
            fa_of_class[c] = (int)fa.size();
            fa.push_back(a);
        }
        // face items in the member's own launch order and partial slots
        for (size_t i = 0; i < s->launch_order.size(); i++) {
            const int c = s->launch_order[i];
            const FaceClass &C = s->P.classes[c];
            const int kind = kind_of(C.ncp, C.npt);
            for (int g = 0; g < face_groups(C.nf()); g++)
                per_member[mi].push_back(FaceItem{fa_of_class[c], g, s->block_start[i] + g, kind});
        }
        GatherArgs g{};
        g.nv = s->T.nv;
        g.vstart = s->d_lvstart;
        g.vrec = s->d_lvrec;
        g.nan_flag = s->d_nan;
        g.gbuf = s->d_gbuf;
        g.gbuf32 = e->mixed ? s->d_gbuf32 : nullptr;
        g.partial = s->d_partial;
        g.npartial = s->npartial;
        g.out_mode = 1;
        g.r = m.d_r;
        g.mobility_dt = m.bd.mobility_dt;
        g.noise_amp = m.bd.noise_amp;
        g.energies = m.d_energies;
        g.energies_cap = m.d_energies ? m.energies_cap : 0;
        g.noise = kNoisePhilox;
        g.seed = m.seed;
        g.step_ctr = s->d_step;
        if (rson) g.rs = s->d_rs;
        const int gi = (int)ga.size();
        ga.push_back(g);
        for (int b = 0; b < gather_blocks(s->T.nv); b++) gitems.push_back(GatherItem{gi, b});
    }
    // items grouped by energy form; within a group, order 0 = member after
    // member, 1 = round-robin over members
    std::vector<FaceItem> items;
    for (int f = 0; f < 3; f++) {
        e->feat_start[f] = (int)items.size();
        std::vector<const std::vector<FaceItem> *> grp;
        for (size_t mi = 0; mi < per_member.size(); mi++)
            if (feat_of[mi] == f) grp.push_back(&per_member[mi]);
        if (e->order == 1) {
            size_t longest = 0;
            for (const auto *v : grp) longest = std::max(longest, v->size());
            for (size_t k = 0; k < longest; k++)
                for (const auto *v : grp)
                    if (k < v->size()) items.push_back((*v)[k]);
        } else {
            for (const auto *v : grp) items.insert(items.end(), v->begin(), v->end());
        }
    }
    e->feat_start[3] = (int)items.size();
    // the restraint finalize's items: the slot-0 item of each restrained
    // member that does not fold the finalize into its face blocks
    std::vector<int> fin;
    for (size_t i = 0; i < items.size(); i++)
        if (items[i].slot == 0 && fa[items[i].fa].rs && !fa[items[i].fa].rs_fold) fin.push_back((int)i);
    int rc;
    if ((rc = upload_table(e, &e->d_fa, fa)) || (rc = upload_table(e, &e->d_items, items)) ||
        (rc = upload_table(e, &e->d_ga, ga)) || (rc = upload_table(e, &e->d_gitems, gitems)) ||
        (rc = upload_table(e, &e->d_fin, fin)))
        return rc;
    e->nfin = (int)fin.size();
    e->nitems = (int)items.size();
    e->ngitems = (int)gitems.size();
    e->nrestrained = nrestrained;
    err = hipStreamSynchronize(e->stream);
    if (err != hipSuccess) return ehip(e, err, "hipStreamSynchronize");
    return UMCP_OK;
}

hipError_t enqueue_step(umcp_ensemble *e, hipStream_t st)
{
    hipError_t err = hipSuccess;
    if (e->nrestrained > 0) {
        // over every item: groups of members without restraints return at once
        BatchFaceArgs rb{e->d_fa, e->d_items, e->nitems, nullptr, nullptr, e->d_fin, e->nfin};
        err = launch_restraint_prepass(rb, st);
        if (err != hipSuccess) return err;
    }
    for (int f = 0; f < 3; f++) {
        const int n = e->feat_start[f + 1] - e->feat_start[f];
        if (n == 0) continue;
        BatchFaceArgs fb{e->d_fa, e->d_items + e->feat_start[f], n, nullptr, nullptr, nullptr, 0};
        err = launch_faces_batch(fb, e->face_grid, e->point_split, e->flags, e->mixed, f, st);
        if (err != hipSuccess) return err;
    }
    BatchGatherArgs gb{e->d_ga, e->d_gitems, e->ngitems};
    return launch_gather_batch(gb, e->gather_grid, e->mixed, st);
}

}  // namespace

extern "C" {

int umcp_ensemble_create(umcp_ensemble **out)
{
    if (!out) return UMCP_ERR_ARG;
    *out = nullptr;
    auto *e = new umcp_ensemble();
    if (hipGetDevice(&e->device) != hipSuccess) {
        delete e;
        return UMCP_ERR_DEVICE;
    }
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
        delete e;
        return UMCP_ERR_DEVICE;
    }
    *out = e;
    return UMCP_OK;
}

void umcp_ensemble_destroy(umcp_ensemble *e)
{
    if (!e) return;
    (void)hipSetDevice(e->device);
    // the caller's stream may be gone by now; hipFree of the tables waits for
    // the device
    (void)hipStreamSynchronize(e->stream);
    drop_graph(e);
    drop_tables(e);
    (void)hipStreamDestroy(e->stream);
    delete e;
}

const char *umcp_ensemble_last_error(const umcp_ensemble *e) { return e ? e->err.c_str() : "null handle"; }

int umcp_ensemble_add(umcp_ensemble *e, umcp_surface *s, double *d_r, const umcp_bd_params *bd,
                      unsigned long long seed, double *d_energies, int energies_cap)
{
    if (!e || !s || !d_r || !bd || energies_cap < 0) return UMCP_ERR_ARG;
    if (!s->planned || !s->g0_set) return efail(e, UMCP_ERR_STATE, "member surface needs umcp_plan and umcp_setg0");
    if (s->device != e->device) return efail(e, UMCP_ERR_ARG, "member surface lives on another device");
    for (const auto &m : e->members)
        if (m.s == s) return efail(e, UMCP_ERR_ARG, "surface is already a member");
    e->members.push_back(umcp_ensemble::Member{s, d_r, *bd, seed, d_energies, energies_cap});
    e->dirty = true;
    return UMCP_OK;
}

int umcp_ensemble_remove(umcp_ensemble *e, umcp_surface *s)
{
    if (!e || !s) return UMCP_ERR_ARG;
    for (size_t i = 0; i < e->members.size(); i++)
        if (e->members[i].s == s) {
            // the device tables still point at the member: finish their work first
            (void)hipSetDevice(e->device);
            hipError_t err = quiesce(e);
            if (err != hipSuccess) return ehip(e, err, "hipStreamSynchronize");
            e->members.erase(e->members.begin() + (long)i);
            e->dirty = true;
            return UMCP_OK;
        }
    return efail(e, UMCP_ERR_ARG, "surface is not a member");
}

int umcp_ensemble_size(const umcp_ensemble *e) { return e ? (int)e->members.size() : UMCP_ERR_ARG; }

int umcp_ensemble_refresh(umcp_ensemble *e)
{
    if (!e) return UMCP_ERR_ARG;
    e->dirty = true;
    return UMCP_OK;
}

int umcp_ensemble_set_launch(umcp_ensemble *e, int point_split, int flags, int face_grid, int order)
{
    if (!e) return UMCP_ERR_ARG;
    if (const char *why = launch_shape_error(point_split, flags)) return efail(e, UMCP_ERR_ARG, why);
    if (face_grid < 0) return efail(e, UMCP_ERR_ARG, "negative grid");
    if (order < 0 || order > 1) return efail(e, UMCP_ERR_ARG, "order is 0 (member-major) or 1 (round-robin)");
    if (order != e->order) e->dirty = true;
    if (point_split != e->point_split || flags != e->flags || face_grid != e->face_grid) drop_graph(e);
    e->point_split = point_split;
    e->flags = flags;
    e->face_grid = face_grid;
    e->order = order;
    return UMCP_OK;
}

int umcp_ensemble_get_launch(const umcp_ensemble *e, int *point_split, int *flags, int *face_grid, int *order)
{
    if (!e) return UMCP_ERR_ARG;
    if (point_split) *point_split = e->point_split;
    if (flags) *flags = e->flags;
    if (face_grid) *face_grid = e->face_grid;
    if (order) *order = e->order;
    return UMCP_OK;
}

int umcp_ensemble_bd_steps(umcp_ensemble *e, int nsteps, int graph_steps, void *stream)
{
    if (!e || nsteps < 0 || graph_steps < 0) return UMCP_ERR_ARG;
    hipError_t err = hipSetDevice(e->device);
    if (err != hipSuccess) return ehip(e, err, "hipSetDevice");
    int rc;
    if (e->dirty && (rc = build(e))) return rc;
    if (e->members.empty() || nsteps == 0) return UMCP_OK;
    hipStream_t st = stream ? (hipStream_t)stream : e->stream;
    if (e->last_stream && e->last_stream != st && e->last_stream != e->stream) {
        // steps on another stream still in flight touch the same positions
        err = hipStreamSynchronize(e->last_stream);
        if (err != hipSuccess) return ehip(e, err, "hipStreamSynchronize");
    }
    e->last_stream = st;
    int done = 0;
    if (graph_steps > 0 && nsteps >= graph_steps) {
        if (!e->exec || e->graph_steps != graph_steps || e->graph_stream != st) {
            drop_graph(e);
            err = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
            if (err != hipSuccess) return ehip(e, err, "hipStreamBeginCapture");
            hipError_t lerr = hipSuccess;
            for (int k = 0; k < graph_steps && lerr == hipSuccess; k++) lerr = enqueue_step(e, st);
            err = hipStreamEndCapture(st, &e->graph);
            if (lerr != hipSuccess) return ehip(e, lerr, "captured launch");
            if (err != hipSuccess) return ehip(e, err, "hipStreamEndCapture");
            err = hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0);
            if (err != hipSuccess) return ehip(e, err, "hipGraphInstantiate");
            e->graph_steps = graph_steps;
            e->graph_stream = st;
        }
        for (; done + graph_steps <= nsteps; done += graph_steps) {
            err = hipGraphLaunch(e->exec, st);
            if (err != hipSuccess) return ehip(e, err, "hipGraphLaunch");
        }
    }
    for (; done < nsteps; done++) {
        err = enqueue_step(e, st);
        if (err != hipSuccess) return ehip(e, err, "batched launch");
    }
    return UMCP_OK;
}

int umcp_ensemble_nan_check(umcp_ensemble *e, int *member)
{
    if (!e) return UMCP_ERR_ARG;
    if (member) *member = -1;
    hipError_t err = hipDeviceSynchronize();
    if (err != hipSuccess) return ehip(e, err, "hipDeviceSynchronize");
    int first = -1;
    for (size_t i = 0; i < e->members.size(); i++) {
        umcp_surface *s = e->members[i].s;
        int flag = 0;
        err = hipMemcpy(&flag, s->d_nan, sizeof flag, hipMemcpyDeviceToHost);
        if (err != hipSuccess) return ehip(e, err, "download nan flag");
        if (!flag) continue;
        if (first < 0) first = (int)i;
        const int zero = 0;
        err = hipMemcpy(s->d_nan, &zero, sizeof zero, hipMemcpyHostToDevice);
        if (err != hipSuccess) return ehip(e, err, "clear nan flag");
    }
    if (first < 0) return UMCP_OK;
    if (member) *member = first;
    return efail(e, UMCP_ERR_NAN_ENERGY, "nan energy in member " + std::to_string(first));
}

int umcp_ensemble_work(const umcp_ensemble *e, int *face_items, int *gather_items, int *restrained)
{
    if (!e) return UMCP_ERR_ARG;
    if (face_items) *face_items = e->nitems;
    if (gather_items) *gather_items = e->ngitems;
    if (restrained) *restrained = e->nrestrained;
    return UMCP_OK;
}

}  // extern "C"
