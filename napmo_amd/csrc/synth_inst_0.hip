// Explicit instantiations of the synthesis kernel for lmax 0..6.
#include "synth_kernel.hpp"
#include "synth_mfma.hpp"

namespace napmo {
template void launch_synth_t<0>(dim3, dim3, hipStream_t, const AtomTables &, const double *,
                                  const double *, const double *, long, const double *,
                                  const double *, const double *, int, int, double *);
template void launch_synth_t<1>(dim3, dim3, hipStream_t, const AtomTables &, const double *,
                                  const double *, const double *, long, const double *,
                                  const double *, const double *, int, int, double *);
template void launch_synth_t<2>(dim3, dim3, hipStream_t, const AtomTables &, const double *,
                                  const double *, const double *, long, const double *,
                                  const double *, const double *, int, int, double *);
template void launch_synth_t<3>(dim3, dim3, hipStream_t, const AtomTables &, const double *,
                                  const double *, const double *, long, const double *,
                                  const double *, const double *, int, int, double *);
template void launch_synth_t<4>(dim3, dim3, hipStream_t, const AtomTables &, const double *,
                                  const double *, const double *, long, const double *,
                                  const double *, const double *, int, int, double *);
template void launch_synth_t<5>(dim3, dim3, hipStream_t, const AtomTables &, const double *,
                                  const double *, const double *, long, const double *,
                                  const double *, const double *, int, int, double *);
template void launch_synth_t<6>(dim3, dim3, hipStream_t, const AtomTables &, const double *,
                                  const double *, const double *, long, const double *,
                                  const double *, const double *, int, int, double *);
template void launch_synth2_t<0>(int, hipStream_t, const AtomTables &, const double *, long, const int *,
                                   const int *, long, const double *, const double *,
                                   const double *, double *);
template void launch_synth2_t<1>(int, hipStream_t, const AtomTables &, const double *, long, const int *,
                                   const int *, long, const double *, const double *,
                                   const double *, double *);
template void launch_synth2_t<2>(int, hipStream_t, const AtomTables &, const double *, long, const int *,
                                   const int *, long, const double *, const double *,
                                   const double *, double *);
template void launch_synth2_t<3>(int, hipStream_t, const AtomTables &, const double *, long, const int *,
                                   const int *, long, const double *, const double *,
                                   const double *, double *);
template void launch_synth2_t<4>(int, hipStream_t, const AtomTables &, const double *, long, const int *,
                                   const int *, long, const double *, const double *,
                                   const double *, double *);
template void launch_synth2_t<5>(int, hipStream_t, const AtomTables &, const double *, long, const int *,
                                   const int *, long, const double *, const double *,
                                   const double *, double *);
template void launch_synth2_t<6>(int, hipStream_t, const AtomTables &, const double *, long, const int *,
                                   const int *, long, const double *, const double *,
                                   const double *, double *);
template void launch_synth_mfma_t<0>(hipStream_t, const AtomTables &, const double *, long,
                                     const int *, const int *, long, int, int, const double *,
                                     const double *, const double *, long, double *);
template void launch_synth_mfma_t<1>(hipStream_t, const AtomTables &, const double *, long,
                                     const int *, const int *, long, int, int, const double *,
                                     const double *, const double *, long, double *);
template void launch_synth_mfma_t<2>(hipStream_t, const AtomTables &, const double *, long,
                                     const int *, const int *, long, int, int, const double *,
                                     const double *, const double *, long, double *);
template void launch_synth_mfma_t<3>(hipStream_t, const AtomTables &, const double *, long,
                                     const int *, const int *, long, int, int, const double *,
                                     const double *, const double *, long, double *);
template void launch_synth_mfma_t<4>(hipStream_t, const AtomTables &, const double *, long,
                                     const int *, const int *, long, int, int, const double *,
                                     const double *, const double *, long, double *);
template void launch_synth_mfma_t<5>(hipStream_t, const AtomTables &, const double *, long,
                                     const int *, const int *, long, int, int, const double *,
                                     const double *, const double *, long, double *);
template void launch_synth_mfma_t<6>(hipStream_t, const AtomTables &, const double *, long,
                                     const int *, const int *, long, int, int, const double *,
                                     const double *, const double *, long, double *);
}  // namespace napmo
