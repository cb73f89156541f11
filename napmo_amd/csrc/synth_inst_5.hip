// Explicit instantiations of the synthesis kernel for lmax 20..21.
#include "synth_kernel.hpp"
#include "synth_mfma.hpp"

namespace napmo {
template void launch_synth_t<20>(dim3, dim3, hipStream_t, const AtomTables &, const double *,
                                  const double *, const double *, long, const double *,
                                  const double *, const double *, int, int, double *);
template void launch_synth_t<21>(dim3, dim3, hipStream_t, const AtomTables &, const double *,
                                  const double *, const double *, long, const double *,
                                  const double *, const double *, int, int, double *);
template void launch_synth2_t<20>(int, hipStream_t, const AtomTables &, const double *, long, const int *,
                                   const int *, long, const double *, const double *,
                                   const double *, double *);
template void launch_synth2_t<21>(int, hipStream_t, const AtomTables &, const double *, long, const int *,
                                   const int *, long, const double *, const double *,
                                   const double *, double *);
template void launch_synth_mfma_t<20>(hipStream_t, const AtomTables &, const double *, long,
                                     const int *, const int *, long, int, int, const double *,
                                     const double *, const double *, long, double *);
template void launch_synth_mfma_t<21>(hipStream_t, const AtomTables &, const double *, long,
                                     const int *, const int *, long, int, int, const double *,
                                     const double *, const double *, long, double *);
}  // namespace napmo
