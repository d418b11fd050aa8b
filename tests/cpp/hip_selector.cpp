// Compile-and-link check of the C++ selector: code written against
// libdress++'s dress::DRESS switches to the GPU by include.
#include "dress/hip/dress.hpp"

#include <cstdio>
#include <vector>

int main()
{
    std::vector<int> U = {0, 1, 2}, V = {1, 2, 0};
    dress::hip::DRESS g(3, U, V);
    dress::DRESS::FitResult f = g.fit(100, 1e-6);
    dress::DRESS::DeltaFitResult d = g.deltaFit(1, 100, 1e-6, 0, 0, true);
    dress::DRESS::NablaFitResult n = g.nablaFit(1, 100, 1e-6);
    auto one = dress::hip::delta_fit(3, U, V, 1);
    std::printf("%d %.17g %zu %lld %zu %lld\n", f.iterations, g.edgeDress(0), d.histogram.size(),
                (long long)d.num_subgraphs, n.histogram.size(), (long long)one.num_subgraphs);
    return 0;
}
