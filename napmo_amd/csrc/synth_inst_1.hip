// Explicit instantiations of the synthesis kernel for lmax 7..11.
#include "synth_kernel.hpp"
#include "synth_mfma.hpp"

namespace napmo {
template void launch_synth_t<7>(dim3, dim3, hipStream_t, const AtomTables &, const double *,
                                  const double *, const double *, long, const double *,
                                  const double *, const double *, int, int, double *);
template void launch_synth_t<8>(dim3, dim3, hipStream_t, const AtomTables &, const double *,
                                  const double *, const double *, long, const double *,
                                  const double *, const double *, int, int, double *);
template void launch_synth_t<9>(dim3, dim3, hipStream_t, const AtomTables &, const double *,
                                  const double *, const double *, long, const double *,
                                  const double *, const double *, int, int, double *);
template void launch_synth_t<10>(dim3, dim3, hipStream_t, const AtomTables &, const double *,
                                  const double *, const double *, long, const double *,
                                  const double *, const double *, int, int, double *);
template void launch_synth_t<11>(dim3, dim3, hipStream_t, const AtomTables &, const double *,
                                  const double *, const double *, long, const double *,
                                  const double *, const double *, int, int, double *);
template void launch_synth2_t<7>(int, hipStream_t, const AtomTables &, const double *, long, const int *,
                                   const int *, long, const double *, const double *,
                                   const double *, double *);
template void launch_synth2_t<8>(int, hipStream_t, const AtomTables &, const double *, long, const int *,
                                   const int *, long, const double *, const double *,
                                   const double *, double *);
template void launch_synth2_t<9>(int, hipStream_t, const AtomTables &, const double *, long, const int *,
                                   const int *, long, const double *, const double *,
                                   const double *, double *);
template void launch_synth2_t<10>(int, hipStream_t, const AtomTables &, const double *, long, const int *,
                                   const int *, long, const double *, const double *,
                                   const double *, double *);
template void launch_synth2_t<11>(int, hipStream_t, const AtomTables &, const double *, long, const int *,
                                   const int *, long, const double *, const double *,
                                   const double *, double *);
template void launch_synth_mfma_t<7>(hipStream_t, const AtomTables &, const double *, long,
                                     const int *, const int *, long, int, int, const double *,
                                     const double *, const double *, long, double *);
template void launch_synth_mfma_t<8>(hipStream_t, const AtomTables &, const double *, long,
                                     const int *, const int *, long, int, int, const double *,
                                     const double *, const double *, long, double *);
template void launch_synth_mfma_t<9>(hipStream_t, const AtomTables &, const double *, long,
                                     const int *, const int *, long, int, int, const double *,
                                     const double *, const double *, long, double *);
template void launch_synth_mfma_t<10>(hipStream_t, const AtomTables &, const double *, long,
                                     const int *, const int *, long, int, int, const double *,
                                     const double *, const double *, long, double *);
template void launch_synth_mfma_t<11>(hipStream_t, const AtomTables &, const double *, long,
                                     const int *, const int *, long, int, int, const double *,
                                     const double *, const double *, long, double *);
}  // namespace napmo
