// Operation count of the oracle's fourth-order box operators (test
// infrastructure, like the oracle itself): hc_oracle.cpp compiled with the
// counting scalar of opcount.hpp in place of double. orc_opcount_box runs the
// standard or Berta box operator on one box and returns the counts of its
// additions/subtractions, multiplications, divisions, square roots and
// comparisons; the geometry helpers do not count (ORC_UNCOUNTED), as the
// device path reads precomputed metric tables. Serial build (no OpenMP).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <limits>
#include <memory>
#include <string>
#include <vector>

#include "opcount.hpp"

#define ORC_UNCOUNTED OpPause orc_uncounted_scope_
#define double Counted
#include "hc_oracle.cpp"
#undef double

extern "C" __attribute__((visibility("default"))) void orc_opcount_box(const orc_cfg* c, int p, const int* lo, const int* hi, const double* U, double* rhs,
                                long long* counts)
{
    opCounts() = OpCounts{};
    opCounting() = true;
    orc_o4_rhs(c, p, lo, hi, reinterpret_cast<const Counted*>(U), reinterpret_cast<Counted*>(rhs),
               nullptr, nullptr, nullptr);
    const OpCounts& k = opCounts();
    counts[0] = k.add; counts[1] = k.mul; counts[2] = k.div; counts[3] = k.sqrt; counts[4] = k.cmp;
    counts[5] = k.other;
}
