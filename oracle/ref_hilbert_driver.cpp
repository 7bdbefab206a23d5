// Thin C-ABI driver around the reference's own hilbert::index
// (include/common/math/hilbert.h:15-48), compiled from the reference tree
// where it lies.  Test infrastructure: used only to pin the oracle's
// restatement of the Hilbert index.
#include <common/math/hilbert.h>

extern "C" long long ref_hilbert_index(int dim, const double* x, long long level) {
  if(dim == 2) {
    VectorD<2> p(x[0], x[1]);
    return hilbert::index<2>(p, level);
  }
  if(dim == 3) {
    VectorD<3> p(x[0], x[1], x[2]);
    return hilbert::index<3>(p, level);
  }
  return -1;
}
