// Points on the limit surface: batched surface::evaluateRNRM
// (src/interpLib.C:7936-8196) and surface::pointGradient (src/grad.C:
// 5697-6443), the calls through which the reference's particle complexes
// couple to the membrane (src/pcomplex.C:501, src/bd.C:1011,
// src/minimize.C:233, src/p_p.C:336-357).
//
// The host resolves each point (face -> class, sub-patch level and local
// coordinates, with the reference's domain rules and rejections) and, for the
// gradient, orders the batch by face (a stable counting sort).  The device
// evaluates one point per thread; the gradient is then summed per (face,
// slot) over the face's points in the caller's order, written into the
// surface's vertex-CSR layout and summed per vertex in CSR order, so results
// do not depend on the launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include "umcp_state.hpp"

using namespace umcp;

namespace {

constexpr double kThresh = 1e-4;   // src/interpLib.C:24

int pfail(umcp_surface *s, int code, const std::string &msg)
{
    s->err = msg;
    return code;
}

int phip(umcp_surface *s, hipError_t e, const char *where)
{
    return pfail(s, UMCP_ERR_DEVICE, std::string(where) + ": " + hipGetErrorString(e));
}

int palloc(umcp_surface *s, void **p, size_t bytes)
{
    hipError_t e = hipMalloc(p, bytes ? bytes : 1);
    if (e != hipSuccess) return pfail(s, UMCP_ERR_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
    s->allocations.push_back(*p);
    s->device_bytes += bytes;
    return UMCP_OK;
}

void pfree(umcp_surface *s, void *p, size_t bytes)
{
    if (!p) return;
    auto it = std::find(s->allocations.begin(), s->allocations.end(), p);
    if (it != s->allocations.end()) s->allocations.erase(it);
    s->device_bytes -= bytes;
    (void)hipFree(p);
}

// grow-only device buffer
int pgrow(umcp_surface *s, char **buf, size_t *cap, size_t need)
{
    if (need <= *cap) return UMCP_OK;
    (void)hipStreamSynchronize(s->stream);
    pfree(s, *buf, *cap);
    *buf = nullptr;
    *cap = 0;
    size_t want = std::max(need, (size_t)1 << 20);
    void *p = nullptr;
    int rc = palloc(s, &p, want);
    if (rc) return rc;
    *buf = (char *)p;
    *cap = want;
    return UMCP_OK;
}

size_t align_up(size_t x) { return (x + kArenaAlign - 1) & ~(kArenaAlign - 1); }

// the per-surface tables: face -> (class, index), the class table and each
// irregular class's sub-patch maps for levels 1..kPointDomains
int point_setup(umcp_surface *s)
{
    if (s->pt_ready) return UMCP_OK;
    const Plan &P = s->P;
    s->pt_face_cls.assign((size_t)P.nfaces, -1);
    s->pt_face_fc.assign((size_t)P.nfaces, -1);
    for (size_t c = 0; c < P.classes.size(); c++)
        for (int i = 0; i < P.classes[c].nf(); i++) {
            s->pt_face_cls[(size_t)P.classes[c].face_id[i]] = (int)c;
            s->pt_face_fc[(size_t)P.classes[c].face_id[i]] = i;
        }
    // one buffer: [class table][maps...]
    std::vector<PointClass> tab(P.classes.size());
    std::vector<size_t> map_off(P.classes.size(), 0);
    const size_t tab_off = 0;
    size_t bytes = align_up(sizeof(PointClass) * tab.size());
    for (size_t c = 0; c < P.classes.size(); c++)
        if (P.classes[c].valence != 6) {
            map_off[c] = bytes;
            bytes += align_up(sizeof(double) * kPointDomains * 3 * 12 * (size_t)P.classes[c].ncp);
        }
    std::vector<char> host(bytes, 0);
    void *dev = nullptr;
    int rc = palloc(s, &dev, bytes);
    if (rc) return rc;
    char *d = (char *)dev;
    for (size_t c = 0; c < P.classes.size(); c++) {
        const FaceClass &C = P.classes[c];
        PointClass &t = tab[c];
        t.cp = s->cls[c].cp;
        t.pos = s->cls[c].pos;
        t.img = s->cls[c].img;
        t.has_img = s->cls[c].has_img;
        t.nf = C.nf();
        t.ncp = C.ncp;
        t.valence = C.valence;
        t.map = nullptr;
        t.g0 = s->cls[c].g0;
        if (C.valence != 6) {
            if (irregular_domain_maps(C.valence, kPointDomains, (double *)(host.data() + map_off[c])) != 0)
                return pfail(s, UMCP_ERR_BAD_MESH, "no sub-patch maps for this valence");
            t.map = (const double *)(d + map_off[c]);
        }
    }
    std::memcpy(host.data() + tab_off, tab.data(), sizeof(PointClass) * tab.size());
    // (face, slot) sums of the gradient in the vertex-CSR layout, and the
    // alpha reduction's partials
    void *gb = nullptr;
    rc = palloc(s, &gb, sizeof(double) * 3 * ((size_t)P.nnz + kPointAlphaParts));
    if (rc) return rc;
    s->d_pt_gbuf = (double *)gb;
    hipError_t e = hipMemcpyAsync(d, host.data(), bytes, hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) return phip(s, e, "upload point tables");
    s->d_pt_cls = (PointClass *)(d + tab_off);
    // the face table of device-resident calls
    std::vector<int2> face((size_t)P.nfaces);
    for (int f = 0; f < P.nfaces; f++) face[(size_t)f] = make_int2(s->pt_face_cls[(size_t)f], s->pt_face_fc[(size_t)f]);
    void *df = nullptr;
    if ((rc = palloc(s, &df, sizeof(int2) * face.size()))) return rc;
    e = hipMemcpyAsync(df, face.data(), sizeof(int2) * face.size(), hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) return phip(s, e, "upload point face table");
    s->d_pt_face = (int2 *)df;
    s->pt_ready = true;
    return UMCP_OK;
}

// irr_kernel::domain / get_map / get_map_transform (src/irr_kernel.C:
// 380-541) on the host, where the level is found with the same libm log10
// the reference uses; the sub-patch arithmetic is exact (powers of two).
// Returns 0 and fills rec, or a message for the inputs the reference stops on.
const char *resolve_irregular(double fu, double fv, bool vertex_ok, PointRec &rec)
{
    double f_domain = -std::log10(fu + fv) / std::log10(2.0);
    if (!(f_domain < 1e300) || !(f_domain > -1e300)) {
        // the extraordinary vertex itself (src/interpLib.C:7963); only the
        // evaluation has a rule there, the reference's map is NULL otherwise
        if (!vertex_ok) return "the extraordinary vertex of an irregular face has no sub-patch (the reference has no map there)";
        if (!(fu < 1e-10 && fv < 1e-10)) return "(u, v) outside the face";
        rec.kind = kPointVertex;
        rec.map = 0;
        rec.u = rec.v = 0.0;
        rec.su = rec.sv = 1.0;
        return nullptr;
    }
    if (fu + fv >= 1.0 - 1e-9) f_domain = 1;
    const int domain = (int)std::lround(std::ceil(f_domain));
    if (domain - 1 >= kPointDomains) return "requested sub-patch level beyond MAX_DOMAIN (100)";
    const double pow2 = std::pow(2.0, domain - 1.0);
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
    if (fu + fv > 1.0 + kThresh || fu < -kThresh || fv < -kThresh) return "(u, v) outside the face";
    rec.kind = kPointIrregular;
    rec.map = (domain - 1) * 3 + sub;
    rec.u = fu;
    rec.v = fv;
    rec.su = su;
    rec.sv = sv;
    return nullptr;
}

struct Batch {
    size_t off_r, off_rec, off_order, off_idx, off_dr, off_dn, off_frac, off_dens, off_seg, in_bytes;
    size_t off_R, off_N, off_gout, off_gp, off_con, off_apart, out_bytes;
};

// pinned grow-only host buffer
int pin_grow(umcp_surface *s, char **buf, size_t *cap, size_t need)
{
    if (need <= *cap) return UMCP_OK;
    (void)hipStreamSynchronize(s->stream);
    if (*buf) (void)hipHostFree(*buf);
    *buf = nullptr;
    *cap = 0;
    const size_t want = std::max(need, (size_t)1 << 20);
    void *p = nullptr;
    hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
    if (e != hipSuccess) return pfail(s, UMCP_ERR_OOM, std::string("hipHostMalloc: ") + hipGetErrorString(e));
    *buf = (char *)p;
    *cap = want;
    return UMCP_OK;
}

// Host threads for a batch of n points: one below 32768 points, else up to
// 16 (the host work is memory bound: records, copies, the face sort).
int host_threads(size_t n)
{
    if (n < 32768) return 1;
    const unsigned hc = std::thread::hardware_concurrency();
    const int t = (int)std::min<unsigned>(hc ? hc : 1u, 16u);
    return std::max(1, std::min(t, (int)(n / 16384)));
}

// f(t, lo, hi) over nt contiguous chunks of [0, n), chunk t on thread t
// (chunk 0 on the caller); a thread that cannot be started runs its chunk
// on the caller instead
template <class F>
void par_chunks(int nt, size_t n, F &&f)
{
    if (nt <= 1) {
        f(0, (size_t)0, n);
        return;
    }
    std::vector<std::thread> th;
    th.reserve((size_t)nt - 1);
    for (int t = 1; t < nt; t++) {
        const size_t lo = n * (size_t)t / (size_t)nt, hi = n * (size_t)(t + 1) / (size_t)nt;
        try {
            th.emplace_back([&f, t, lo, hi] { f(t, lo, hi); });
        } catch (const std::system_error &) {
            f(t, lo, hi);
        }
    }
    f(0, (size_t)0, n / (size_t)nt);
    for (auto &x : th) x.join();
}

void par_copy(int nt, void *dst, const void *src, size_t bytes)
{
    par_chunks(nt, bytes, [&](int, size_t lo, size_t hi) {
        std::memcpy((char *)dst + lo, (const char *)src + lo, hi - lo);
    });
}

// resolves the points, orders a gradient batch by face, stages and launches;
// mode 0 evaluation, 1 gradient, 2 geometry, 3 control-point weights.
// Records go straight into the pinned staging buffer, over host threads in
// contiguous chunks of the batch; a rejection names the first bad point, as
// a serial pass would.
// mode 0 evaluation, 1 point gradient, 2 geometry, 3 weights, 4 particle
// gradient (de_dr = de_dnrm: kc a, c0, 0 per particle; face_pfa and face_c0
// per face in global order: the particle area of grad's density term and the
// face's spontaneous curvature)
int run_points(umcp_surface *s, const double *r, int n, const int *faces, const double *uv, const double *de_dr,
               const double *de_dnrm, const double *frac, int on_surface, double kT, int mode, Batch &b,
               const double *face_pfa = nullptr, const double *face_c0 = nullptr)
{
    const bool grad = mode == 1 || mode == 4;
    const int nv = s->T.nv;
    const Plan &P = s->P;
    const int nfaces = P.nfaces;
    // staging layout; the segment table (at most one per face with points)
    // goes last so that only its used part is uploaded
    const size_t seg_cap = grad ? std::min<size_t>((size_t)n, (size_t)nfaces) : 0;
    size_t o = 0;
    b.off_r = o;      o = align_up(o + sizeof(double) * (3 * (size_t)nv + 3));
    b.off_rec = o;    o = align_up(o + sizeof(PointRec) * (size_t)n);
    // the face-sorted positions run past n when faces get a density record
    const size_t order_cap = grad ? (size_t)n + (mode == 4 ? seg_cap : 0) : 0;
    b.off_order = o;  o = align_up(o + sizeof(int) * order_cap);
    b.off_idx = o;    o = align_up(o + sizeof(int) * (size_t)n);
    b.off_dr = o;     o = align_up(o + (grad ? sizeof(double) * 3 * (size_t)n : 0));
    b.off_dn = o;     o = align_up(o + (grad ? sizeof(double) * 3 * (size_t)n : 0));
    b.off_frac = o;   o = align_up(o + (grad && frac ? sizeof(double) * (size_t)n : 0));
    b.off_dens = o;   o = align_up(o + (mode == 4 ? sizeof(DensityItem) * seg_cap : 0));
    b.off_seg = o;    o = align_up(o + sizeof(PointSeg) * seg_cap);
    const size_t in_cap = o;
    int rc = pin_grow(s, &s->pt_stage, &s->pt_stage_cap, in_cap);
    if (rc) return rc;
    char *h = s->pt_stage;
    PointRec *rec = (PointRec *)(h + b.off_rec);
    int *order = (int *)(h + b.off_order), *idx = (int *)(h + b.off_idx);
    PointSeg *seg = (PointSeg *)(h + b.off_seg);

    const int nt = host_threads((size_t)n);
    struct Chunk {
        int counts[3] = {0, 0, 0};
        int bad = -1;              // first rejected point of the chunk
        const char *msg = nullptr;
        std::vector<int> fcount;   // gradient: points per face in the chunk
    };
    std::vector<Chunk> ch((size_t)nt);
    try {
        // allocated here, not in the workers, where a failure could not be reported
        if (grad)
            for (Chunk &c : ch) c.fcount.assign((size_t)nfaces, 0);
    } catch (const std::bad_alloc &) {
        return pfail(s, UMCP_ERR_OOM, "host memory for the face sort");
    }
    const std::vector<int> &fcls = s->pt_face_cls, &ffc = s->pt_face_fc;
    par_chunks(nt, (size_t)n, [&](int t, size_t lo, size_t hi) {
        Chunk &c = ch[(size_t)t];
        for (size_t i = lo; i < hi; i++) {
            const int f = faces[i];
            if (f < 0 || f >= nfaces) {
                c.bad = (int)i;
                c.msg = "face index out of range";
                return;
            }
            PointRec &q = rec[i];
            q.cls = fcls[(size_t)f];
            q.fc = ffc[(size_t)f];
            const double u = uv[2 * i], v = uv[2 * i + 1];
            q.fu = u;
            q.fv = v;
            if (!std::isfinite(u) || !std::isfinite(v)) {
                c.bad = (int)i;
                c.msg = "non-finite (u, v)";
                return;
            }
            if (P.classes[(size_t)q.cls].valence == 6) {
                q.kind = kPointRegular;
                q.map = 0;
                q.u = u;
                q.v = v;
                q.su = q.sv = 1.0;
            } else if (const char *m = resolve_irregular(u, v, mode == 0 || mode == 3, q)) {
                c.bad = (int)i;
                c.msg = m;
                return;
            }
            c.counts[q.kind]++;
            if (grad) c.fcount[(size_t)f]++;
        }
    });
    for (const Chunk &c : ch)
        if (c.bad >= 0) return pfail(s, UMCP_ERR_ARG, "point " + std::to_string(c.bad) + ": " + c.msg);
    int counts[3] = {0, 0, 0};
    for (const Chunk &c : ch)
        for (int k = 0; k < 3; k++) counts[k] += c.counts[k];

    int nseg = 0, ndens = 0;
    DensityItem *dens = (DensityItem *)(h + b.off_dens);
    if (grad) {
        // stable counting sort by face: chunk t's points of face f start at
        // the face's start plus the earlier chunks' counts of f, so the order
        // is that of one serial pass.  One segment per face with points; the
        // kind is the face's (regular or irregular; the vertex itself is
        // rejected above), so the kind lists are runs of segments.
        int fill[3] = {0, counts[0], counts[0] + counts[1]};
        int pos = 0;
        for (int f = 0; f < nfaces; f++) {
            int cnt = 0;
            for (Chunk &c : ch) {
                const int k = c.fcount[(size_t)f];
                c.fcount[(size_t)f] = pos + cnt;
                cnt += k;
            }
            if (!cnt) continue;
            // particle gradient: a face carrying particle area in grad's
            // density term gets one more record after its particles
            const int extra = (mode == 4 && face_pfa[(size_t)f] != 0.0) ? 1 : 0;
            seg[nseg++] = {fcls[(size_t)f], ffc[(size_t)f], pos, pos + cnt + extra};
            if (extra)
                dens[ndens++] = {fcls[(size_t)f], ffc[(size_t)f], pos + cnt, 0, face_pfa[(size_t)f],
                                 face_c0[(size_t)f]};
            const int kind = P.classes[(size_t)fcls[(size_t)f]].valence == 6 ? kPointRegular : kPointIrregular;
            for (int j = pos; j < pos + cnt; j++) idx[fill[kind]++] = j;
            pos += cnt + extra;
        }
        par_chunks(nt, (size_t)n, [&](int t, size_t lo, size_t hi) {
            std::vector<int> &next = ch[(size_t)t].fcount;
            for (size_t i = lo; i < hi; i++) order[next[(size_t)faces[i]]++] = (int)i;
            std::memcpy(h + b.off_dr + sizeof(double) * 3 * lo, de_dr + 3 * lo, sizeof(double) * 3 * (hi - lo));
            std::memcpy(h + b.off_dn + sizeof(double) * 3 * lo, de_dnrm + 3 * lo, sizeof(double) * 3 * (hi - lo));
            if (frac) std::memcpy(h + b.off_frac + sizeof(double) * lo, frac + lo, sizeof(double) * (hi - lo));
        });
    } else {
        // positions grouped by kind, each group in the caller's order
        std::vector<int> base((size_t)nt * 3);
        int fill[3] = {0, counts[0], counts[0] + counts[1]};
        for (int t = 0; t < nt; t++)
            for (int k = 0; k < 3; k++) {
                base[(size_t)t * 3 + k] = fill[k];
                fill[k] += ch[(size_t)t].counts[k];
            }
        par_chunks(nt, (size_t)n, [&](int t, size_t lo, size_t hi) {
            int *nx = &base[(size_t)t * 3];
            for (size_t i = lo; i < hi; i++) idx[nx[rec[i].kind]++] = (int)i;
        });
    }
    if (r) std::memcpy(h + b.off_r, r, sizeof(double) * (3 * (size_t)nv + 3));
    else std::memset(h + b.off_r, 0, sizeof(double) * (3 * (size_t)nv + 3));
    b.in_bytes = b.off_seg + sizeof(PointSeg) * (size_t)nseg;
    const int nrec = n + ndens;   // gradient records per slot
    o = 0;
    if (grad) {
        // [vertex sums 3 nv][alpha 3][gpuv 2n] first: one download
        b.off_gout = o;
        b.off_gp = o + sizeof(double) * (3 * (size_t)nv + 3);
        o = align_up(b.off_gp + sizeof(double) * 2 * (size_t)n);
        b.off_con = o;    o = align_up(o + sizeof(double) * (size_t)nrec * kPointSlots * 3);
        b.off_apart = o;  o = align_up(o + sizeof(double) * (size_t)nseg * kPointSlots * 3);
        b.off_R = b.off_N = 0;
    } else if (mode == 2) {
        b.off_R = o;   // the geometry records
        o = align_up(o + sizeof(double) * kPointGeom * (size_t)n);
        b.off_N = b.off_gout = b.off_gp = b.off_con = b.off_apart = 0;
    } else if (mode == 3) {
        b.off_R = o;   // weights, then vertex ids
        b.off_N = align_up(o + sizeof(double) * 3 * kPointSlots * (size_t)n);
        o = align_up(b.off_N + sizeof(int) * kPointSlots * (size_t)n);
        b.off_gout = b.off_gp = b.off_con = b.off_apart = 0;
    } else {
        b.off_R = o;
        b.off_N = o + sizeof(double) * 3 * (size_t)n;
        o = align_up(b.off_N + sizeof(double) * 3 * (size_t)n);
        b.off_gout = b.off_gp = b.off_con = b.off_apart = 0;
    }
    b.out_bytes = o;
    rc = pgrow(s, &s->d_pt_in, &s->pt_in_cap, b.in_bytes);
    if (rc == UMCP_OK) rc = pgrow(s, &s->d_pt_out, &s->pt_out_cap, b.out_bytes);
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(s->d_pt_in, h, b.in_bytes, hipMemcpyHostToDevice, s->stream);
    if (e != hipSuccess) return phip(s, e, "upload points");
    PointArgs a{};
    char *di = s->d_pt_in, *dout = s->d_pt_out;
    a.r = (const double *)(di + b.off_r);
    a.nv = nv;
    a.n = nrec;
    a.grad = mode;
    a.on_surface = on_surface ? 1 : 0;
    a.kT = kT;
    a.kc = s->prm.kc;
    std::memcpy(a.pbc, s->T.pbc, sizeof a.pbc);
    a.rec = (const PointRec *)(di + b.off_rec);
    a.cls = s->d_pt_cls;
    if (grad) {
        a.order = (const int *)(di + b.off_order);
        a.de_dr = (const double *)(di + b.off_dr);
        a.de_dnrm = (const double *)(di + b.off_dn);
        a.frac = frac ? (const double *)(di + b.off_frac) : nullptr;
        a.contrib = (double *)(dout + b.off_con);
        a.gpuv = (double *)(dout + b.off_gp);
    } else if (mode == 2) {
        a.geom = (double *)(dout + b.off_R);
    } else if (mode == 3) {
        a.wts = (double *)(dout + b.off_R);
        a.ids = (int *)(dout + b.off_N);
    } else {
        a.R = (double *)(dout + b.off_R);
        a.nrm = (double *)(dout + b.off_N);
    }
    e = launch_points(a, (const int *)(di + b.off_idx), counts, s->stream);
    if (e != hipSuccess) return phip(s, e, "point kernel launch");
    if (mode == 4) {
        e = launch_particle_density(a, (const DensityItem *)(di + b.off_dens), ndens, s->stream);
        if (e != hipSuccess) return phip(s, e, "particle density kernel launch");
    }
    if (grad) {
        PointReduceArgs ra{};
        ra.r = a.r;
        ra.nv = nv;
        ra.n = nrec;
        ra.nseg = nseg;
        ra.nnz = P.nnz;
        std::memcpy(ra.pbc, s->T.pbc, sizeof ra.pbc);
        ra.seg = (const PointSeg *)(di + b.off_seg);
        ra.cls = s->d_pt_cls;
        ra.contrib = a.contrib;
        ra.vstart = s->d_vstart;
        ra.gbuf = s->d_pt_gbuf;
        ra.apart = (double *)(dout + b.off_apart);
        ra.partial = s->d_pt_gbuf + 3 * (size_t)P.nnz;
        ra.gout = (double *)(dout + b.off_gout);
        e = launch_point_reduce(ra, s->stream);
        if (e != hipSuccess) return phip(s, e, "point reduce launch");
    }
    return UMCP_OK;
}

// copies [off, off + bytes) of the batch's device output into pinned host
// memory (the caller's arrays may be pageable) and returns it there
int fetch_out(umcp_surface *s, size_t off, size_t bytes, const char **host, const char *what)
{
    int rc = pin_grow(s, &s->pt_host, &s->pt_host_cap, bytes);
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(s->pt_host, s->d_pt_out + off, bytes, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) return phip(s, e, what);
    *host = s->pt_host;
    return UMCP_OK;
}

int point_check(umcp_surface *s, const double *r, int n, const int *faces, const double *uv, bool need_r = true)
{
    if (!s) return UMCP_ERR_ARG;
    if (!s->planned) return pfail(s, UMCP_ERR_STATE, "plan has not been generated (call umcp_plan)");
    if (n < 0 || (n > 0 && ((need_r && !r) || !faces || !uv))) return pfail(s, UMCP_ERR_ARG, "null point arrays");
    if (n > UMCP_POINT_MAX_BATCH) return pfail(s, UMCP_ERR_ARG, "more points than UMCP_POINT_MAX_BATCH");
    hipError_t e = hipSetDevice(s->device);
    if (e != hipSuccess) return phip(s, e, "hipSetDevice");
    return point_setup(s);
}

}  // namespace

extern "C" {

int umcp_point_eval(umcp_surface *s, const double *r, int n, const int *faces, const double *uv, double *R,
                    double *nrm)
{
    int rc = point_check(s, r, n, faces, uv);
    if (rc) return rc;
    if (n == 0) return UMCP_OK;
    if (!R || !nrm) return pfail(s, UMCP_ERR_ARG, "null output arrays");
    Batch b;
    rc = run_points(s, r, n, faces, uv, nullptr, nullptr, nullptr, 0, 0.0, 0, b);
    if (rc) return rc;
    const size_t bytes = sizeof(double) * 3 * (size_t)n;   // R, then nrm
    const char *hst = nullptr;
    rc = fetch_out(s, b.off_R, b.off_N - b.off_R + bytes, &hst, "download points");
    if (rc) return rc;
    const int nt = host_threads((size_t)n);
    par_copy(nt, R, hst, bytes);
    par_copy(nt, nrm, hst + (b.off_N - b.off_R), bytes);
    return UMCP_OK;
}

// the device-resident evaluation (mode 0: R, nrm) or geometry (mode 2: geom)
static int point_dev_run(umcp_surface *s, const double *d_r, int n, const int *d_faces, const double *d_uv, int mode,
                         double *d_R, double *d_nrm, double *d_geom, void *stream)
{
    int rc = point_check(s, d_r, n, d_faces, d_uv);
    if (rc) return rc;
    if (n > 0 && (mode == 0 ? (!d_R || !d_nrm) : !d_geom)) return pfail(s, UMCP_ERR_ARG, "null output arrays");
    hipStream_t st = stream ? (hipStream_t)stream : s->stream;
    // work space: records, keys and indices (twice: sort in and out), the
    // per-kind counts, the sort's temporary storage; the status word apart
    size_t temp = 0;
    hipError_t e = point_partition_by_kind(nullptr, nullptr, nullptr, nullptr, std::max(n, 1), nullptr, &temp, st);
    if (e != hipSuccess) return phip(s, e, "point partition size");
    size_t o = 0;
    const size_t o_rec = o;   o = align_up(o + sizeof(PointRec) * (size_t)n);
    const size_t o_k0 = o;    o = align_up(o + (size_t)n);
    const size_t o_k1 = o;    o = align_up(o + (size_t)n);
    const size_t o_v0 = o;    o = align_up(o + sizeof(int) * (size_t)n);
    const size_t o_v1 = o;    o = align_up(o + sizeof(int) * (size_t)n);
    const size_t o_cnt = o;   o = align_up(o + sizeof(int) * 3);
    const size_t o_tmp = o;   o = align_up(o + temp);
    if (o > s->ptd_cap) {
        if (s->d_ptd) {
            // earlier calls may still use it, on any stream
            (void)hipDeviceSynchronize();
            pfree(s, s->d_ptd, s->ptd_cap);
            s->d_ptd = nullptr;
            s->ptd_cap = 0;
        }
        void *p = nullptr;
        if ((rc = palloc(s, &p, o))) return rc;
        s->d_ptd = (char *)p;
        s->ptd_cap = o;
    }
    if (!s->d_ptd_status) {
        void *p = nullptr;
        if ((rc = palloc(s, &p, sizeof(unsigned long long)))) return rc;
        s->d_ptd_status = (unsigned long long *)p;
    }
    char *w = s->d_ptd;
    PointResolveArgs ra{};
    ra.n = n;
    ra.nfaces = s->P.nfaces;
    ra.vertex_ok = mode == 0 ? 1 : 0;   // only the evaluation has a rule at the vertex
    ra.faces = d_faces;
    ra.uv = d_uv;
    ra.face = s->d_pt_face;
    ra.cls = s->d_pt_cls;
    ra.rec = (PointRec *)(w + o_rec);
    ra.kind = (unsigned char *)(w + o_k0);
    ra.iota = (int *)(w + o_v0);
    ra.counts = (int *)(w + o_cnt);
    ra.status = s->d_ptd_status;
    ra.R = mode == 0 ? d_R : nullptr;
    ra.nrm = mode == 0 ? d_nrm : nullptr;
    ra.geom = mode == 2 ? d_geom : nullptr;
    e = launch_point_resolve(ra, st);
    if (e != hipSuccess) return phip(s, e, "point resolve launch");
    if (n == 0) return UMCP_OK;
    e = point_partition_by_kind(ra.kind, (unsigned char *)(w + o_k1), ra.iota, (int *)(w + o_v1), n, w + o_tmp, &temp,
                                st);
    if (e == hipSuccess) e = launch_point_counts((const unsigned char *)(w + o_k1), n, ra.counts, st);
    if (e != hipSuccess) return phip(s, e, "point partition");
    PointArgs a{};
    a.r = d_r;
    a.nv = s->T.nv;
    a.n = n;
    a.grad = mode;
    a.kc = s->prm.kc;
    std::memcpy(a.pbc, s->T.pbc, sizeof a.pbc);
    a.rec = ra.rec;
    a.cls = s->d_pt_cls;
    a.R = mode == 0 ? d_R : nullptr;
    a.nrm = mode == 0 ? d_nrm : nullptr;
    a.geom = mode == 2 ? d_geom : nullptr;
    e = launch_points_dev(a, (const int *)(w + o_v1), ra.counts, n, st);
    if (e != hipSuccess) return phip(s, e, "point kernel launch");
    return UMCP_OK;
}

int umcp_point_eval_dev(umcp_surface *s, const double *d_r, int n, const int *d_faces, const double *d_uv,
                        double *d_R, double *d_nrm, void *stream)
{
    return point_dev_run(s, d_r, n, d_faces, d_uv, 0, d_R, d_nrm, nullptr, stream);
}

int umcp_point_geometry_dev(umcp_surface *s, const double *d_r, int n, const int *d_faces, const double *d_uv,
                            double *d_geom, void *stream)
{
    return point_dev_run(s, d_r, n, d_faces, d_uv, 2, nullptr, nullptr, d_geom, stream);
}

int umcp_point_grad_dev(umcp_surface *s, const double *d_r, int n, const int *d_faces, const double *d_uv,
                        const double *d_de_dr, const double *d_de_dnrm, const double *d_frac_mult, int on_surface,
                        double kT, double *d_g, double *d_g_puv, void *stream)
{
    int rc = point_check(s, d_r, n, d_faces, d_uv);
    if (rc) return rc;
    if (n > 0 && (!d_de_dr || !d_de_dnrm || !d_g || !d_g_puv)) return pfail(s, UMCP_ERR_ARG, "null gradient arrays");
    hipStream_t st = stream ? (hipStream_t)stream : s->stream;
    const Plan &P = s->P;
    const int nv = s->T.nv, nfaces = P.nfaces;
    const int nseg_max = std::max(1, std::min(n, nfaces));
    int fbits = 1;
    while ((1 << fbits) <= nfaces) fbits++;   // keys up to nfaces (rejected points)
    const int nn = std::max(n, 1);
    size_t t_face = 0, t_kind = 0, t_scan = 0;
    hipError_t e = point_sort_by_face(nullptr, nullptr, nullptr, nullptr, nn, fbits, nullptr, &t_face, st);
    if (e == hipSuccess) e = point_partition_by_kind(nullptr, nullptr, nullptr, nullptr, nn, nullptr, &t_kind, st);
    if (e == hipSuccess) e = point_inclusive_sum(nullptr, nullptr, nn, nullptr, &t_scan, st);
    if (e != hipSuccess) return phip(s, e, "point sort sizes");
    size_t temp = std::max(t_face, std::max(t_kind, t_scan));
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o = align_up(o + bytes);
        return at;
    };
    const size_t o_rec = take(sizeof(PointRec) * (size_t)n), o_kind = take((size_t)n),
                 o_iota = take(sizeof(int) * (size_t)n), o_order = take(sizeof(int) * (size_t)n),
                 o_fkey = take(sizeof(int) * (size_t)n), o_skey = take(sizeof(int) * (size_t)n),
                 o_flag = take(sizeof(int) * (size_t)n), o_incl = take(sizeof(int) * (size_t)n),
                 o_pk0 = take((size_t)n), o_pk1 = take((size_t)n), o_idx = take(sizeof(int) * (size_t)n),
                 o_cnt = take(sizeof(int) * 3), o_seg = take(sizeof(PointSeg) * (size_t)nseg_max),
                 o_con = take(sizeof(double) * 3 * kPointSlots * (size_t)n), o_gp = take(sizeof(double) * 2 * (size_t)n),
                 o_ap = take(sizeof(double) * 3 * kPointSlots * (size_t)nseg_max),
                 o_gout = take(sizeof(double) * (3 * (size_t)nv + 3)), o_tmp = take(temp);
    if (o > s->ptd_cap) {
        if (s->d_ptd) {
            // earlier calls may still use it, on any stream
            (void)hipDeviceSynchronize();
            pfree(s, s->d_ptd, s->ptd_cap);
            s->d_ptd = nullptr;
            s->ptd_cap = 0;
        }
        void *p = nullptr;
        if ((rc = palloc(s, &p, o))) return rc;
        s->d_ptd = (char *)p;
        s->ptd_cap = o;
    }
    if (!s->d_ptd_status) {
        void *p = nullptr;
        if ((rc = palloc(s, &p, sizeof(unsigned long long)))) return rc;
        s->d_ptd_status = (unsigned long long *)p;
    }
    char *w = s->d_ptd;
    PointResolveArgs ra{};
    ra.n = n;
    ra.nfaces = nfaces;
    ra.vertex_ok = 0;   // the gradient has no rule at the extraordinary vertex
    ra.faces = d_faces;
    ra.uv = d_uv;
    ra.face = s->d_pt_face;
    ra.cls = s->d_pt_cls;
    ra.rec = (PointRec *)(w + o_rec);
    ra.kind = (unsigned char *)(w + o_kind);
    ra.iota = (int *)(w + o_iota);
    ra.counts = (int *)(w + o_cnt);
    ra.status = s->d_ptd_status;
    ra.fkey = (int *)(w + o_fkey);
    e = launch_point_resolve(ra, st);
    if (e != hipSuccess) return phip(s, e, "point resolve launch");
    if (n == 0) return UMCP_OK;
    // positions in face order (stable: a face's points keep the caller's
    // order), the faces' segments, the positions grouped by kind
    int *order = (int *)(w + o_order), *skey = (int *)(w + o_skey), *flags = (int *)(w + o_flag),
        *incl = (int *)(w + o_incl), *idx = (int *)(w + o_idx);
    unsigned char *pk0 = (unsigned char *)(w + o_pk0), *pk1 = (unsigned char *)(w + o_pk1);
    PointSeg *seg = (PointSeg *)(w + o_seg);
    size_t tb = temp;
    e = point_sort_by_face(ra.fkey, skey, ra.iota, order, n, fbits, w + o_tmp, &tb, st);
    if (e == hipSuccess) e = launch_point_seg_flags(skey, n, nfaces, flags, st);
    tb = temp;
    if (e == hipSuccess) e = point_inclusive_sum(flags, incl, n, w + o_tmp, &tb, st);
    if (e == hipSuccess) e = launch_point_seg_fill(skey, incl, n, nfaces, s->d_pt_face, s->d_pt_cls, seg, pk0, st);
    tb = temp;
    if (e == hipSuccess) e = point_partition_by_kind(pk0, pk1, ra.iota, idx, n, w + o_tmp, &tb, st);
    if (e == hipSuccess) e = launch_point_counts(pk1, n, ra.counts, st);
    if (e != hipSuccess) return phip(s, e, "point face sort");
    PointArgs a{};
    a.r = d_r;
    a.nv = nv;
    a.n = n;
    a.grad = 1;
    a.on_surface = on_surface ? 1 : 0;
    a.kT = kT;
    a.kc = s->prm.kc;
    std::memcpy(a.pbc, s->T.pbc, sizeof a.pbc);
    a.rec = ra.rec;
    a.cls = s->d_pt_cls;
    a.order = order;
    a.de_dr = d_de_dr;
    a.de_dnrm = d_de_dnrm;
    a.frac = d_frac_mult;
    a.contrib = (double *)(w + o_con);
    a.gpuv = (double *)(w + o_gp);
    e = launch_points_dev(a, idx, ra.counts, n, st);
    if (e != hipSuccess) return phip(s, e, "point kernel launch");
    PointReduceArgs rd{};
    rd.r = d_r;
    rd.nv = nv;
    rd.n = n;
    rd.nseg = 0;
    rd.nnz = P.nnz;
    std::memcpy(rd.pbc, s->T.pbc, sizeof rd.pbc);
    rd.seg = seg;
    rd.cls = s->d_pt_cls;
    rd.contrib = a.contrib;
    rd.vstart = s->d_vstart;
    rd.gbuf = s->d_pt_gbuf;
    rd.apart = (double *)(w + o_ap);
    rd.partial = s->d_pt_gbuf + 3 * (size_t)P.nnz;
    rd.gout = (double *)(w + o_gout);
    rd.dnseg = incl + (n - 1);
    e = launch_point_reduce_dev(rd, nseg_max, st);
    // the reference's contract: the caller's arrays accumulate (+=)
    if (e == hipSuccess) e = launch_point_accumulate(rd.gout, d_g, 3 * nv + 3, 1, nullptr, st);
    if (e == hipSuccess) e = launch_point_accumulate(a.gpuv, d_g_puv, n, 2, ra.kind, st);
    if (e != hipSuccess) return phip(s, e, "point reduce launch");
    return UMCP_OK;
}

int umcp_point_status(umcp_surface *s, int *bad_point)
{
    if (!s) return UMCP_ERR_ARG;
    if (bad_point) *bad_point = -1;
    if (!s->d_ptd_status) return UMCP_OK;
    unsigned long long st = 0;
    hipError_t e = hipMemcpy(&st, s->d_ptd_status, sizeof st, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return phip(s, e, "point status");
    if (st == ~0ull) return UMCP_OK;
    static const char *why[] = {"", "face index out of range", "non-finite (u, v)",
                                "the extraordinary vertex of an irregular face has no sub-patch (the reference has no map there)",
                                "(u, v) outside the face", "requested sub-patch level beyond MAX_DOMAIN (100)"};
    const int i = (int)(st >> 8), c = (int)(st & 0xff);
    if (bad_point) *bad_point = i;
    return pfail(s, UMCP_ERR_ARG, "point " + std::to_string(i) + ": " + (c >= 1 && c <= 5 ? why[c] : "rejected"));
}

int umcp_point_grad(umcp_surface *s, const double *r, int n, const int *faces, const double *uv,
                    const double *de_dr, const double *de_dnrm, const double *frac_mult, int on_surface, double kT,
                    double *g, double *g_puv)
{
    int rc = point_check(s, r, n, faces, uv);
    if (rc) return rc;
    if (n == 0) return UMCP_OK;
    if (!de_dr || !de_dnrm || !g || !g_puv) return pfail(s, UMCP_ERR_ARG, "null gradient arrays");
    Batch b;
    rc = run_points(s, r, n, faces, uv, de_dr, de_dnrm, frac_mult, on_surface, kT, 1, b);
    if (rc) return rc;
    const size_t ng = 3 * (size_t)s->T.nv + 3;   // vertex sums, alpha, then gpuv: contiguous
    const char *hst = nullptr;
    rc = fetch_out(s, b.off_gout, sizeof(double) * (ng + 2 * (size_t)n), &hst, "download point gradient");
    if (rc) return rc;
    const double *gs = (const double *)hst, *gp = gs + ng;
    // the reference's contract: the caller's arrays accumulate (+=)
    for (size_t i = 0; i < ng; i++) g[i] += gs[i];
    par_chunks(host_threads((size_t)n), 2 * (size_t)n, [&](int, size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; i++) g_puv[i] += gp[i];
    });
    return UMCP_OK;
}

int umcp_particle_grad(umcp_surface *s, const double *r, int n, const int *faces, const double *uv,
                       const double *c0, const double *area, double *g, double *g_puv)
{
    int rc = point_check(s, r, n, faces, uv);
    if (rc) return rc;
    if (n == 0) return UMCP_OK;
    if (!c0 || !area || !g || !g_puv) return pfail(s, UMCP_ERR_ARG, "null particle arrays");
    const int nfaces = s->P.nfaces;
    // grad's density term counts the particles whose c0 differs from their
    // face's by 1e-7 or more (src/grad.C:93-103, 6483-6487)
    std::vector<double> pfa((size_t)nfaces, 0.0), cf((size_t)nfaces, 0.0), pc(3 * (size_t)n, 0.0);
    if (!s->face_c0.empty()) cf = s->face_c0;
    for (int i = 0; i < n; i++) {
        const int f = faces[i];
        if (f < 0 || f >= nfaces) return pfail(s, UMCP_ERR_ARG, "point " + std::to_string(i) + ": face index out of range");
        if (std::fabs(c0[i] - cf[(size_t)f]) >= 1e-7) pfa[(size_t)f] += area[i];
        pc[3 * (size_t)i] = s->prm.kc * area[i];
        pc[3 * (size_t)i + 1] = c0[i];
    }
    Batch b;
    rc = run_points(s, r, n, faces, uv, pc.data(), pc.data(), nullptr, 0, 0.0, 4, b, pfa.data(), cf.data());
    if (rc) return rc;
    const size_t ng = 3 * (size_t)s->T.nv + 3;
    const char *hst = nullptr;
    rc = fetch_out(s, b.off_gout, sizeof(double) * (ng + 2 * (size_t)n), &hst, "download particle gradient");
    if (rc) return rc;
    const double *gs = (const double *)hst, *gp = gs + ng;
    for (size_t i = 0; i < ng; i++) g[i] += gs[i];
    for (size_t i = 0; i < 2 * (size_t)n; i++) g_puv[i] += gp[i];
    return UMCP_OK;
}

int umcp_point_coeffs(umcp_surface *s, int n, const int *faces, const double *uv, int *ids, double *w)
{
    int rc = point_check(s, nullptr, n, faces, uv, false);
    if (rc) return rc;
    if (n == 0) return UMCP_OK;
    if (!ids || !w) return pfail(s, UMCP_ERR_ARG, "null output arrays");
    Batch b;
    rc = run_points(s, nullptr, n, faces, uv, nullptr, nullptr, nullptr, 0, 0.0, 3, b);
    if (rc) return rc;
    const size_t wb = sizeof(double) * 3 * kPointSlots * (size_t)n, ib = sizeof(int) * kPointSlots * (size_t)n;
    const char *hst = nullptr;
    rc = fetch_out(s, b.off_R, b.off_N - b.off_R + ib, &hst, "download point weights");
    if (rc) return rc;
    const int nt = host_threads((size_t)n);
    par_copy(nt, w, hst, wb);
    par_copy(nt, ids, hst + (b.off_N - b.off_R), ib);
    // at the extraordinary vertex the reference lists the ring in the
    // vertex's own (oriented) edge order, src/uv_map.C:46-52; the face's
    // control-point table holds the same vertices rotated to the face's edge
    const PointRec *rec = (const PointRec *)(s->pt_stage + b.off_rec);
    for (int i = 0; i < n; i++) {
        if (rec[i].kind != kPointVertex) continue;
        int *id = ids + (size_t)i * kPointSlots;
        const int v = id[0], val = s->T.valence[v];
        for (int k = 0; k < val; k++) id[1 + k] = s->T.ring[(size_t)v * kMaxValence + k];
    }
    return UMCP_OK;
}

int umcp_particle_energy(umcp_surface *s, const double *r, int n, const int *faces, const double *uv,
                         const double *c0, const double *area, double *out)
{
    int rc = point_check(s, r, n, faces, uv);
    if (rc) return rc;
    if (n == 0) return UMCP_OK;
    if (!c0 || !area || !out) return pfail(s, UMCP_ERR_ARG, "null particle arrays");
    const int nf = s->P.nfaces;
    std::vector<double> num((size_t)nf), a0((size_t)nf);
    if ((rc = umcp::face_density(s, r, num.data(), a0.data()))) return rc;
    Batch b;
    rc = run_points(s, r, n, faces, uv, nullptr, nullptr, nullptr, 0, 0.0, 2, b);
    if (rc) return rc;
    const char *hst = nullptr;
    rc = fetch_out(s, b.off_R, sizeof(double) * kPointGeom * (size_t)n, &hst, "download particle curvatures");
    if (rc) return rc;
    const double *G = (const double *)hst;
    const double kc = s->prm.kc;
    // the reference's order: en += 1/2 kc a (c1 + c2 - c0)^2; en -= a fed with
    // fed = face_energy_density / face_area (src/interpLib.C:8580-8584, 3636)
    par_chunks(host_threads((size_t)n), (size_t)n, [&](int, size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; i++) {
            const double c1 = G[i * kPointGeom + 9], c2 = G[i * kPointGeom + 10];
            const int f = faces[i];
            const double fed = num[(size_t)f] / a0[(size_t)f];
            double en = 0.0;
            en += 0.5 * kc * area[i] * (c1 + c2 - c0[i]) * (c1 + c2 - c0[i]);
            en -= area[i] * fed;
            out[i] = en;
        }
    });
    return UMCP_OK;
}

int umcp_point_geometry(umcp_surface *s, const double *r, int n, const int *faces, const double *uv, double *geom)
{
    int rc = point_check(s, r, n, faces, uv);
    if (rc) return rc;
    if (n == 0) return UMCP_OK;
    if (!geom) return pfail(s, UMCP_ERR_ARG, "null output array");
    Batch b;
    rc = run_points(s, r, n, faces, uv, nullptr, nullptr, nullptr, 0, 0.0, 2, b);
    if (rc) return rc;
    const size_t bytes = sizeof(double) * kPointGeom * (size_t)n;
    const char *hst = nullptr;
    rc = fetch_out(s, b.off_R, bytes, &hst, "download point geometry");
    if (rc) return rc;
    par_copy(host_threads((size_t)n), geom, hst, bytes);
    return UMCP_OK;
}

}  // extern "C"

namespace {

// face corners (slot 0 at (u, v) = (0, 0), slot 2 at (1, 0), slot 1 at
// (0, 1)) and the face across each edge, built once
int walk_setup(umcp_surface *s)
{
    if (s->nf_ready) return UMCP_OK;
    const Plan &P = s->P;
    s->nf_corner.assign(3 * (size_t)P.nfaces, -1);
    for (const FaceClass &C : P.classes)
        for (int i = 0; i < C.nf(); i++) {
            const int f = C.face_id[(size_t)i];
            s->nf_corner[3 * (size_t)f] = C.cp[(size_t)0 * C.nf() + i];
            s->nf_corner[3 * (size_t)f + 1] = C.cp[(size_t)2 * C.nf() + i];
            s->nf_corner[3 * (size_t)f + 2] = C.cp[(size_t)1 * C.nf() + i];
        }
    std::map<std::pair<int, int>, std::vector<int>> edges;
    for (int f = 0; f < P.nfaces; f++)
        for (int k = 0; k < 3; k++) {
            int a = s->nf_corner[3 * (size_t)f + k], b = s->nf_corner[3 * (size_t)f + (k + 1) % 3];
            edges[{std::min(a, b), std::max(a, b)}].push_back(f);
        }
    s->nf_across.assign(3 * (size_t)P.nfaces, -1);
    for (int f = 0; f < P.nfaces; f++) {
        // edge 0: v = 0 (corners 0, 1); edge 1: u + v = 1 (1, 2); edge 2: u = 0 (0, 2)
        const int pairs[3][2] = {{0, 1}, {1, 2}, {0, 2}};
        for (int k = 0; k < 3; k++) {
            int a = s->nf_corner[3 * (size_t)f + pairs[k][0]], b = s->nf_corner[3 * (size_t)f + pairs[k][1]];
            const auto &lst = edges[{std::min(a, b), std::max(a, b)}];
            for (int g : lst)
                if (g != f) s->nf_across[3 * (size_t)f + k] = g;
            if (lst.size() != 2 || s->nf_across[3 * (size_t)f + k] < 0)
                return pfail(s, UMCP_ERR_BAD_MESH, "an edge is not shared by exactly two faces");
        }
    }
    s->nf_ready = true;
    return UMCP_OK;
}

// (u, v) of vertex a among the corners of face f
bool corner_uv(const umcp_surface *s, int f, int a, double uv[2])
{
    const double at[3][2] = {{0.0, 0.0}, {1.0, 0.0}, {0.0, 1.0}};
    for (int k = 0; k < 3; k++)
        if (s->nf_corner[3 * (size_t)f + k] == a) {
            uv[0] = at[k][0];
            uv[1] = at[k][1];
            return true;
        }
    return false;
}

}  // namespace

extern "C" {

int umcp_next_face(umcp_surface *s, const double *r, int n, int *faces, double *uv, double *duv, double *mom,
                   double *coord_transform)
{
    int rc = point_check(s, r, n, faces, uv);
    if (rc) return rc;
    if (n == 0) return UMCP_OK;
    if (!duv) return pfail(s, UMCP_ERR_ARG, "null displacement array");
    if ((rc = walk_setup(s))) return rc;
    const int nfaces = s->P.nfaces;
    // per crossing point: the old face at the crossing and the new face there
    std::vector<int> cross_idx, gf;
    std::vector<double> guv, rem;
    for (int i = 0; i < n; i++) {
        const int f = faces[i];
        if (f < 0 || f >= nfaces) return pfail(s, UMCP_ERR_ARG, "point " + std::to_string(i) + ": face index out of range");
        double u = uv[2 * i], v = uv[2 * i + 1], du = duv[2 * i], dv = duv[2 * i + 1];
        if (du == 0 && dv == 0) continue;
        if (u + v > 1 + kThresh || u < -kThresh || v < -kThresh)
            return pfail(s, UMCP_ERR_ARG, "point " + std::to_string(i) + ": not on its face");
        const double nu = u + du, nv = v + dv;
        if (nu > 0 && nv > 0 && nu + nv < 1) {
            uv[2 * i] += du;
            uv[2 * i + 1] += dv;
            duv[2 * i] = duv[2 * i + 1] = 0.0;
            continue;
        }
        // moving along an edge: the reference's nudges (src/uv_map.C:1743-1749)
        if (std::fabs(du + dv) < 1e-10 && std::fabs(du) + std::fabs(dv) > 1e-10) du += 1e-12;
        if (std::fabs(du) < 1e-10 && std::fabs(dv) > 1e-10) du += 1e-12;
        if (std::fabs(dv) < 1e-10 && std::fabs(du) > 1e-10) du += 1e-12;
        double t1 = 1e10, t2 = 1e10, t3 = 1e10;
        if (dv < 0) t1 = -v / dv;
        if (du + dv > 0) t2 = (1 - u - v) / (du + dv);
        if (du < 0) t3 = -u / du;
        int edge;
        double t;
        if (t1 < t2 && t1 < t3) { edge = 0; t = t1; }
        else if (t2 < t3) { edge = 1; t = t2; }
        else { edge = 2; t = t3; }
        u += t * du;
        v += t * dv;
        du -= t * du;
        dv -= t * dv;
        // the shared edge a -> b and the crossing's parameter along it
        const int c0 = s->nf_corner[3 * (size_t)f], c1 = s->nf_corner[3 * (size_t)f + 1],
                  c2 = s->nf_corner[3 * (size_t)f + 2];
        int a, b;
        double sp;
        if (edge == 0) { a = c0; b = c1; sp = u; }
        else if (edge == 1) { a = c1; b = c2; sp = v; }
        else { a = c0; b = c2; sp = v; }
        const int g = s->nf_across[3 * (size_t)f + edge];
        double ua[2], ub[2];
        if (!corner_uv(s, g, a, ua) || !corner_uv(s, g, b, ub))
            return pfail(s, UMCP_ERR_BAD_MESH, "inconsistent face corners");
        const double nu2 = (1 - sp) * ua[0] + sp * ub[0], nv2 = (1 - sp) * ua[1] + sp * ub[1];
        cross_idx.push_back(i);
        gf.push_back(f);
        gf.push_back(g);
        guv.insert(guv.end(), {u, v, nu2, nv2});
        rem.insert(rem.end(), {du, dv});
    }
    const int m = (int)cross_idx.size();
    if (m == 0) return UMCP_OK;
    // tangents at the crossings, on both sides (surface::ru / rv)
    std::vector<double> geom((size_t)2 * m * kPointGeom);
    Batch b;
    rc = run_points(s, r, 2 * m, gf.data(), guv.data(), nullptr, nullptr, nullptr, 0, 0.0, 2, b);
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(geom.data(), s->d_pt_out + b.off_R, sizeof(double) * geom.size(),
                                  hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) return phip(s, e, "download tangents");
    for (int x = 0; x < m; x++) {
        const int i = cross_idx[(size_t)x];
        const double *go = &geom[(size_t)(2 * x) * kPointGeom], *gn = &geom[(size_t)(2 * x + 1) * kPointGeom];
        const double *ruo = go, *rvo = go + 3, *run = gn, *rvn = gn + 3;
        auto dot = [](const double *p, const double *q) { return p[0] * q[0] + p[1] * q[1] + p[2] * q[2]; };
        // (du r_u + dv r_v = du' r_u' + dv' r_v') projected on r_u and r_v of
        // the old face (src/uv_map.C:1876-1893)
        const double ruo1 = dot(ruo, ruo), run1 = dot(run, ruo), ruo2 = dot(ruo, rvo), run2 = dot(run, rvo);
        const double rvo1 = dot(rvo, ruo), rvn1 = dot(rvn, ruo), rvo2 = dot(rvo, rvo), rvn2 = dot(rvn, rvo);
        const double du = rem[2 * (size_t)x], dv = rem[2 * (size_t)x + 1];
        const double den = run2 * rvn1 - run1 * rvn2;
        const double dup = -(-(du * ruo2 * rvn1) + du * ruo1 * rvn2 + dv * rvn2 * rvo1 - dv * rvn1 * rvo2) / den;
        const double dvp = -(-(du * run2 * ruo1) + du * run1 * ruo2 - dv * run2 * rvo1 + dv * run1 * rvo2) / den;
        if (mom) {
            const double rot[4] = {(ruo2 * rvn1 - ruo1 * rvn2) / den, -(rvn2 * rvo1 - rvn1 * rvo2) / den,
                                   (run2 * ruo1 - run1 * ruo2) / den, (run2 * rvo1 - run1 * rvo2) / den};
            if (coord_transform) {
                double *ct = coord_transform + 4 * (size_t)i;
                const double tt[4] = {rot[0] * ct[0] + rot[1] * ct[2], rot[0] * ct[1] + rot[1] * ct[3],
                                      rot[2] * ct[0] + rot[3] * ct[2], rot[2] * ct[1] + rot[3] * ct[3]};
                std::memcpy(ct, tt, sizeof tt);
            }
            const double mu = mom[2 * i] * rot[0] + mom[2 * i + 1] * rot[1];
            const double mv = mom[2 * i] * rot[2] + mom[2 * i + 1] * rot[3];
            mom[2 * i] = mu;
            mom[2 * i + 1] = mv;
        }
        faces[i] = gf[2 * (size_t)x + 1];
        uv[2 * i] = guv[4 * (size_t)x + 2];
        uv[2 * i + 1] = guv[4 * (size_t)x + 3];
        duv[2 * i] = dup;
        duv[2 * i + 1] = dvp;
    }
    return UMCP_OK;
}

}  // extern "C"
