// Explicit instantiations of the synthesis kernel for lmax 22..23.
#include "synth_kernel.hpp"

namespace napmo {
template void launch_synth_t<22>(dim3, dim3, hipStream_t, const AtomTables &, const double *,
                                  const double *, const double *, long, const double *,
                                  const double *, const double *, int, int, double *);
template void launch_synth_t<23>(dim3, dim3, hipStream_t, const AtomTables &, const double *,
                                  const double *, const double *, long, const double *,
                                  const double *, const double *, int, int, double *);
template void launch_synth2_t<22>(int, hipStream_t, const AtomTables &, const double *, long, const int *,
                                   const int *, long, const double *, const double *,
                                   const double *, double *);
template void launch_synth2_t<23>(int, hipStream_t, const AtomTables &, const double *, long, const int *,
                                   const int *, long, const double *, const double *,
                                   const double *, double *);
}  // namespace napmo
