// afd_srhd.hip -- SRHD face-kernel instantiations (MUSCL/PPM/WENO x HLLE/HLLC x x,y,z x FP64/mixed).
#include "afd_launch.hpp"

namespace afd {

template <int RECON, int RIEM, typename T> static void go(const FaceLaunch &l, hipStream_t s) {
    FaceArgs<T> a = face_args<T>(l);
    dim3 grid = grid1d(face_count(l.g, l.dir));
    if (l.dir == 0) hipLaunchKernelGGL((k_face_srhd<RECON, RIEM, 0, T>), grid, dim3(BLOCK), 0, s, a);
    else if (l.dir == 1) hipLaunchKernelGGL((k_face_srhd<RECON, RIEM, 1, T>), grid, dim3(BLOCK), 0, s, a);
    else hipLaunchKernelGGL((k_face_srhd<RECON, RIEM, 2, T>), grid, dim3(BLOCK), 0, s, a);
}

template <typename T> static void go_t(const FaceLaunch &l, hipStream_t s) {
    int key = l.recon * 2 + (l.riemann == 1 ? 1 : 0);
    switch (key) {
        case 0: go<0, 0, T>(l, s); break;
        case 1: go<0, 1, T>(l, s); break;
        case 2: go<1, 0, T>(l, s); break;
        case 3: go<1, 1, T>(l, s); break;
        case 4: go<2, 0, T>(l, s); break;
        default: go<2, 1, T>(l, s); break;
    }
}

hipError_t launch_face_srhd(const FaceLaunch &l, hipStream_t s) {
    if (l.mixed) go_t<float>(l, s);
    else go_t<double>(l, s);
    return hipGetLastError();
}

}  // namespace afd
