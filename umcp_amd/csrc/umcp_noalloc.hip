// umcp_energy_grad_dev_noalloc: the energy and gradient of umcp_energy_grad_dev
// with every work buffer of the evaluation (the per-class face tables, the
// (face, slot) gradient contributions, the block partials and the restraint
// partials and scalars) carved from one file-scope device array instead of
// the handle's own buffers, for callers that cannot allocate (code under
// stream capture, serving loops that forbid allocation).  The kernels are the
// same as umcp_energy_grad_dev's and run in the same order, so the results
// are bitwise equal.
//
// The array is `scratch`, kScratchBytes per device.  Host code never passes
// its name to a kernel: the device address is looked up once per device with
// hipGetSymbolAddress (device current) and kept in g_scratch.  Every call on
// a device shares that device's array; calls that overlap on one device race.
// Nothing needs to start at zero: each call writes every entry it reads
// before reading it (one writer per (face, slot) record and per partial).
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/umcp.h"
#include "umcp_kernels.hpp"
#include "umcp_state.hpp"

namespace {

constexpr size_t kScratchBytes = size_t(8) << 20;   // UMCP_NOALLOC_SCRATCH_BYTES
static_assert(kScratchBytes == UMCP_NOALLOC_SCRATCH_BYTES, "header and library disagree on the scratch size");
constexpr size_t kAlign = 256;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void *g_scratch[kMaxDevices] = {};

int fail(umcp_surface *s, int code, const std::string &msg)
{
    s->err = msg;
    return code;
}

// device address of this device's scratch array (looked up on first use)
int scratch_ptr(umcp_surface *s, char **out)
{
    if (s->device < 0 || s->device >= kMaxDevices) return fail(s, UMCP_ERR_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[s->device]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(s, UMCP_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[s->device] = p;
    }
    *out = (char *)g_scratch[s->device];
    return UMCP_OK;
}

size_t carve(size_t &off, size_t bytes)
{
    const size_t at = (off + kAlign - 1) & ~(kAlign - 1);
    off = at + bytes;
    return at;
}

}  // namespace

extern "C" int umcp_energy_grad_dev_noalloc(umcp_surface *s, const double *d_r, double *d_E, double *d_g,
                                            void *stream)
{
    using namespace umcp;
    if (!s || !d_r || !d_E) return s ? fail(s, UMCP_ERR_ARG, "null argument") : UMCP_ERR_ARG;
    if (!s->planned) return fail(s, UMCP_ERR_STATE, "plan has not been generated (call umcp_plan)");
    if (!s->g0_set) return fail(s, UMCP_ERR_STATE, "reference state not set (call umcp_setg0)");
    hipError_t e = hipSetDevice(s->device);
    if (e != hipSuccess) return fail(s, UMCP_ERR_DEVICE, std::string("hipSetDevice: ") + hipGetErrorString(e));
    const bool mixed = s->prm.precision == UMCP_PREC_MIXED;
    const bool rson = restraints_on(s);
    const size_t ncls = s->P.classes.size();
    const size_t ngroups = (size_t)s->npartial;

    // layout of this call's buffers in the scratch array
    size_t off = 0;
    const size_t o_fa = carve(off, ncls * sizeof(FaceArgs));
    const size_t o_partial = carve(off, 4 * ngroups * sizeof(double));
    const size_t o_rs = carve(off, 8 * sizeof(double));
    const size_t o_rpart = carve(off, 2 * ngroups * sizeof(double));
    const size_t o_gbuf = carve(off, 3 * (size_t)s->nrecords * (mixed ? sizeof(float) : sizeof(double)));
    if (off > kScratchBytes)
        return fail(s, UMCP_ERR_ARG,
                    "surface needs " + std::to_string(off) + " bytes of scratch; the no-allocation entry point holds " +
                        std::to_string(kScratchBytes));
    char *base = nullptr;
    if (int rc = scratch_ptr(s, &base)) return rc;
    FaceArgs *d_fa = (FaceArgs *)(base + o_fa);
    double *partial = (double *)(base + o_partial);
    double *rs = (double *)(base + o_rs);
    double *rpart = (double *)(base + o_rpart);

    std::vector<FaceArgs> &fa = s->h_fa_noalloc;
    fa.resize(ncls);
    for (size_t c = 0; c < ncls; c++) {
        fa[c] = face_args(s, c, d_r);
        fa[c].gbuf = mixed ? nullptr : (double *)(base + o_gbuf);
        fa[c].gbuf32 = mixed ? (float *)(base + o_gbuf) : nullptr;
        fa[c].partial = partial;
        fa[c].rs = rson ? rs : nullptr;
        fa[c].rpart = rpart;
        fa[c].step_ctr = nullptr;
    }
    hipStream_t st = stream ? (hipStream_t)stream : s->stream;
    e = hipMemcpyAsync(d_fa, fa.data(), ncls * sizeof(FaceArgs), hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return fail(s, UMCP_ERR_DEVICE, std::string("upload face table: ") + hipGetErrorString(e));

    BatchFaceArgs fb{};
    fb.fa = d_fa;
    fb.items = s->d_items;
    fb.nitems = s->npartial;
    fb.r = d_r;
    fb.step_ctr = nullptr;
    fb.fin = s->d_fin;
    fb.nfin = s->npartial > kFoldMaxGroups ? 1 : 0;   // else folded into the face kernel
    if (rson) {
        e = launch_restraint_prepass(fb, st);
        if (e != hipSuccess) return fail(s, UMCP_ERR_DEVICE, std::string("restraint kernel launch: ") + hipGetErrorString(e));
    }
    e = launch_faces_batch(fb, 0, s->point_split, s->launch_flags, mixed, features_of(fa[0].ph), st);
    if (e != hipSuccess) return fail(s, UMCP_ERR_DEVICE, std::string("face kernel launch: ") + hipGetErrorString(e));
    GatherArgs ga{};
    ga.out_mode = d_g ? 0 : 2;
    ga.g = d_g;
    ga.E_out = d_E;
    ga.nv = s->T.nv;
    ga.vstart = s->d_lvstart;
    ga.vrec = s->d_lvrec;
    ga.nan_flag = s->d_nan;
    ga.gbuf = mixed ? nullptr : (const double *)(base + o_gbuf);
    ga.gbuf32 = mixed ? (const float *)(base + o_gbuf) : nullptr;
    ga.partial = partial;
    ga.npartial = s->npartial;
    ga.rs = rson ? rs : nullptr;
    e = launch_vertex_gather(ga, st);
    if (e != hipSuccess) return fail(s, UMCP_ERR_DEVICE, std::string("gather kernel launch: ") + hipGetErrorString(e));
    return UMCP_OK;
}
