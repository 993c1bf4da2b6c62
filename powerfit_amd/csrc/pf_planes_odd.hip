// Plane kernels of the run-time planned fused pipeline for odd x lengths
// (pf_planes_kern.h, ODD = true), in a translation unit of their own so the
// two forms compile in parallel.  Rows y = 2p and 2p + 1 travel as one
// complex row through an nx-point FFT and are separated, or packed, by
// Hermitian symmetry.
#include "pf_planes_kern.h"

namespace pf {

int prepare_plane_kernels_odd() { return prepare_plane_kernels<true>(); }

int launch_plane_kernel_odd(const Fft3D& f, const FusedArgs& a, int which, bool skip, int zr,
                            hipStream_t s) {
    return launch_plane_kernel<true>(f, a, which, skip, zr, s);
}

}  // namespace pf
