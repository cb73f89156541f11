// Test infrastructure only: C-linkage shims over reference functions that the
// reference builds without C linkage. Compiled together with the reference's own
// sources (where they lie) into oracle/_ref/libnapmo_ref.so by oracle/Makefile.
// Nothing in the product links this.
void ld_by_order(int order, double *x, double *y, double *z, double *w);
long fill_pure_polynomials(double *output, long lmax);

extern "C" {
void ref_ld_by_order(int order, double *x, double *y, double *z, double *w) {
  ld_by_order(order, x, y, z, w);
}
long ref_fill_pure_polynomials(double *output, long lmax) {
  return fill_pure_polynomials(output, lmax);
}
}
