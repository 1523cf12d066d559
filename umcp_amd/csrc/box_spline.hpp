// The 12 quartic box-spline basis functions of a regular Loop patch
// (src/interpLib.C:5836-5918; slot permutation ceff_map, 6183-6189) as
// monomial tables, shared by the host plan and the device point kernels.
#pragma once

namespace umcp {

// basis n1..n12 as {a, b, coefficient} of u^a v^b w^(4-a-b), times 12
constexpr int kBoxTerms[12][16][3] = {
    {{4, 0, 1}, {3, 1, 2}},
    {{4, 0, 1}, {3, 0, 2}},
    {{4, 0, 1}, {3, 0, 2}, {3, 1, 6}, {2, 1, 6}, {2, 2, 12}, {1, 2, 6}, {1, 3, 6}, {0, 3, 2}, {0, 4, 1}},
    {{4, 0, 6}, {3, 0, 24}, {2, 0, 24}, {1, 0, 8}, {0, 0, 1}, {3, 1, 24}, {2, 1, 60}, {1, 1, 36},
     {0, 1, 6}, {2, 2, 24}, {1, 2, 36}, {0, 2, 12}, {1, 3, 8}, {0, 3, 6}, {0, 4, 1}},
    {{4, 0, 1}, {3, 0, 6}, {2, 0, 12}, {1, 0, 6}, {0, 0, 1}, {3, 1, 2}, {2, 1, 6}, {1, 1, 6}, {0, 1, 2}},
    {{1, 3, 2}, {0, 4, 1}},
    {{4, 0, 1}, {3, 0, 6}, {2, 0, 12}, {1, 0, 6}, {0, 0, 1}, {3, 1, 8}, {2, 1, 36}, {1, 1, 36},
     {0, 1, 8}, {2, 2, 24}, {1, 2, 60}, {0, 2, 24}, {1, 3, 24}, {0, 3, 24}, {0, 4, 6}},
    {{4, 0, 1}, {3, 0, 8}, {2, 0, 24}, {1, 0, 24}, {0, 0, 6}, {3, 1, 6}, {2, 1, 36}, {1, 1, 60},
     {0, 1, 24}, {2, 2, 12}, {1, 2, 36}, {0, 2, 24}, {1, 3, 6}, {0, 3, 8}, {0, 4, 1}},
    {{1, 0, 2}, {0, 0, 1}},
    {{0, 3, 2}, {0, 4, 1}},
    {{1, 0, 2}, {0, 0, 1}, {1, 1, 6}, {0, 1, 6}, {1, 2, 6}, {0, 2, 12}, {1, 3, 2}, {0, 3, 6}, {0, 4, 1}},
    {{0, 0, 1}, {0, 1, 2}},
};
// control-point slot y carries basis n_(kSlotBasis[y])
constexpr int kSlotBasis[12] = {8, 7, 4, 5, 9, 12, 11, 10, 6, 3, 1, 2};

// the 15 quartic monomials u^a v^b w^(4-a-b), a ascending, then b ascending
constexpr int kMono = 15;
constexpr int mono_index(int a, int b) { return a * (11 - a) / 2 + b; }

// monomial coefficients of basis n_(n+1), times 12
struct BoxTable {
    double c[12][kMono];
};
constexpr BoxTable make_box_table()
{
    BoxTable b{};
    for (int n = 0; n < 12; n++)
        for (int t = 0; t < 16 && kBoxTerms[n][t][2]; t++)
            b.c[n][mono_index(kBoxTerms[n][t][0], kBoxTerms[n][t][1])] += kBoxTerms[n][t][2];
    return b;
}
constexpr BoxTable kBoxTable = make_box_table();

}  // namespace umcp
