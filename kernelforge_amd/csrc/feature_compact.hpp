// feature_compact.hpp -- drop descriptor columns that are zero for every atom of the sets
// a call works on (feature_compact.hip).
#pragma once

#include "kf_common.hpp"

namespace kf {

// One or two sets of descriptors X (nm, ma, rep) with optional Jacobians dX
// (nm, ma, rep, 3 ma).  Set 2 may be the same arrays as set 1 (then it is gathered once).
struct CompactSets {
    const double *x1 = nullptr, *dx1 = nullptr, *x2 = nullptr, *dx2 = nullptr;
    const int *n1 = nullptr, *n2 = nullptr;
    int nm1 = 0, ma1 = 0, nm2 = 0, ma2 = 0;
};

// Find the columns that are non-zero in X or dX for some atom (i < N[a]) of either set; if
// some column is unused, gather both sets onto the used ones (in column order) into the
// device workspace slots ws1 / ws2 and point `s` at the copies, with *nf their width.
// When every column is used, `s` is left as it is and *nf = rep.  Synchronises `stream`.
int compact_columns(CompactSets &s, int rep, int ws1, int ws2, hipStream_t stream, int *nf);

}  // namespace kf
