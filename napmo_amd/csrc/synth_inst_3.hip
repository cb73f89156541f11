// Explicit instantiations of the synthesis kernel for lmax 15..17.
#include "synth_kernel.hpp"
#include "synth_mfma.hpp"

namespace napmo {
template void launch_synth_t<15>(dim3, dim3, hipStream_t, const AtomTables &, const double *,
                                  const double *, const double *, long, const double *,
                                  const double *, const double *, int, int, double *);
template void launch_synth_t<16>(dim3, dim3, hipStream_t, const AtomTables &, const double *,
                                  const double *, const double *, long, const double *,
                                  const double *, const double *, int, int, double *);
template void launch_synth_t<17>(dim3, dim3, hipStream_t, const AtomTables &, const double *,
                                  const double *, const double *, long, const double *,
                                  const double *, const double *, int, int, double *);
template void launch_synth2_t<15>(int, hipStream_t, const AtomTables &, const double *, long, const int *,
                                   const int *, long, const double *, const double *,
                                   const double *, double *);
template void launch_synth2_t<16>(int, hipStream_t, const AtomTables &, const double *, long, const int *,
                                   const int *, long, const double *, const double *,
                                   const double *, double *);
template void launch_synth2_t<17>(int, hipStream_t, const AtomTables &, const double *, long, const int *,
                                   const int *, long, const double *, const double *,
                                   const double *, double *);
template void launch_synth_mfma_t<15>(hipStream_t, const AtomTables &, const double *, long,
                                     const int *, const int *, long, int, int, const double *,
                                     const double *, const double *, long, double *);
template void launch_synth_mfma_t<16>(hipStream_t, const AtomTables &, const double *, long,
                                     const int *, const int *, long, int, int, const double *,
                                     const double *, const double *, long, double *);
template void launch_synth_mfma_t<17>(hipStream_t, const AtomTables &, const double *, long,
                                     const int *, const int *, long, int, int, const double *,
                                     const double *, const double *, long, double *);
}  // namespace napmo
