// elem::Print / elem::Write on a small matrix: prints the Print output to
// stdout, then writes the Write output to the file named by argv[1].  Host
// code only; no device call is made.
#include <cstdio>

#include "elemental_amd/elem.hpp"

int main(int argc, char** argv) {
    elem::Matrix<double> A(2, 3);
    const double v[2][3] = {{1.0, 2.5, -0.125}, {-3.0, 1e-7, 123456789.0}};
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 3; ++j) A.Set(i, j, v[i][j]);
    elem::Print(A, "A");
    elem::Matrix<float> B(1, 2);
    B.Set(0, 0, 0.5f);
    B.Set(0, 1, -7.25f);
    elem::Print(B);
    if (argc > 1) elem::Write(A, "A written", argv[1]);
    return 0;
}
