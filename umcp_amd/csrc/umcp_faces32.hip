// FP32-mixed face kernels (faces_batch_kernel<float, ...>) in every launch shape
// and energy form; compiled as their own translation unit.
#include "face_device.hpp"

namespace umcp {

template <int WS, int OCC, int FEAT>
static hipError_t launch_ws32(const BatchFaceArgs &a, int nb, bool persist, hipStream_t st)
{
    if (persist)
        hipLaunchKernelGGL((faces_batch_kernel<float, WS, true, OCC, FEAT, true>), dim3(nb), dim3(kWave * WS), 0, st,
                           a);
    else
        hipLaunchKernelGGL((faces_batch_kernel<float, WS, true, OCC, FEAT, false>), dim3(nb), dim3(kWave * WS), 0,
                           st, a);
    return hipGetLastError();
}

// the control-point split (CPS form) of the kernel
template <int WS, int FEAT>
static hipError_t launch_cps32(const BatchFaceArgs &a, int nb, bool persist, hipStream_t st)
{
    if (persist)
        hipLaunchKernelGGL((faces_batch_kernel<float, WS, true, 1, FEAT, true, true>), dim3(nb), dim3(kWave * WS), 0,
                           st, a);
    else
        hipLaunchKernelGGL((faces_batch_kernel<float, WS, true, 1, FEAT, false, true>), dim3(nb), dim3(kWave * WS), 0,
                           st, a);
    return hipGetLastError();
}

// the register-bound flag asks for 4 waves per SIMD (not in the CPS form)
template <int FEAT>
static hipError_t launch_shape32(const BatchFaceArgs &a, int nb, bool persist, int ws, int flags, hipStream_t st)
{
    if (flags & kLaunchCps) {
        // the control-point split needs its registers only (no bound)
        switch (ws) {
        case 2: return launch_cps32<2, FEAT>(a, nb, persist, st);
        case 3: return launch_cps32<3, FEAT>(a, nb, persist, st);
        case 4: return launch_cps32<4, FEAT>(a, nb, persist, st);
        default: return hipErrorInvalidValue;
        }
    }
    const int occ2 = (flags & kLaunchOcc2) ? 1 : 0;
    switch (ws * 2 + occ2) {
    case 2: return launch_ws32<1, 1, FEAT>(a, nb, persist, st);
    case 3: return launch_ws32<1, 4, FEAT>(a, nb, persist, st);
    case 4: return launch_ws32<2, 1, FEAT>(a, nb, persist, st);
    case 5: return launch_ws32<2, 4, FEAT>(a, nb, persist, st);
    case 6: return launch_ws32<3, 1, FEAT>(a, nb, persist, st);
    case 7: return launch_ws32<3, 4, FEAT>(a, nb, persist, st);
    case 8: return launch_ws32<4, 1, FEAT>(a, nb, persist, st);
    case 9: return launch_ws32<4, 4, FEAT>(a, nb, persist, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_faces32(const BatchFaceArgs &a, int nb, bool persist, int ws, int flags, int feat, hipStream_t st)
{
    switch (feat) {
    case kFeatLipid: return launch_shape32<kFeatLipid>(a, nb, persist, ws, flags, st);
    case kFeatHelfrich: return launch_shape32<kFeatHelfrich>(a, nb, persist, ws, flags, st);
    default: return launch_shape32<kFeatGeneral>(a, nb, persist, ws, flags, st);
    }
}

hipError_t copy_tables_faces32(int kind, const double *buf, const float *b32, const double *corner, size_t n, int npt,
                               int ncp)
{
    return copy_tables_here(kind, buf, b32, corner, n, npt, ncp);
}

}  // namespace umcp
