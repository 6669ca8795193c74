// get_or_compute of the C++ mirror (include/sparkx_hb.hpp) against a cache directory
// (argv[1]): the first call runs HyperBall and saves registers and accumulators, a
// computed centrality is written to its file, the second call reopens all of it.
// Prints the results of both so the test can compare them with the oracle.
#include <cstdio>
#include <vector>

#include "../../include/sparkx_hb.hpp"

int main(int argc, char **argv)
{
    if (argc < 2) return 3;
    std::vector<uint64_t> off = {0, 1, 3, 5, 7, 8, 10, 12, 14};
    std::vector<uint64_t> nbr = {1, 0, 2, 1, 3, 2, 4, 3, 6, 7, 5, 7, 5, 6};
    sparkx::CsrView g{off.data(), 8, 8, nbr.data(), 8, nbr.size()};
    try {
        std::vector<double> c1;
        {
            auto hb = sparkx::HyperBall6::get_or_compute(g, argv[1], "cppgraph");
            c1 = hb.compute_closeness_centrality(true);
            printf("first iters %zu\n", hb.get_iters());
        }
        auto hb = sparkx::HyperBall6::get_or_compute(g, argv[1], "cppgraph");
        printf("second iters %zu\n", hb.get_iters());
        const auto &c2 = hb.get_closeness_centrality(true);  // found in the cache
        const auto dist = hb.geodesic_distances();
        const auto &lin = hb.compute_lins_centrality();      // epilogue on the reopened run
        for (size_t u = 0; u < dist.size(); ++u)
            printf("node %zu %.17g %.17g %.17g %.17g\n", u, dist[u], c1[u], c2[u], lin[u]);
        try {
            (void)hb.get_hamonic_centrality(false);
            printf("error: harmonic was never computed\n");
            return 1;
        } catch (const sparkx::HyperBallError &) {
        }
    } catch (const sparkx::HyperBallError &e) {
        printf("HyperBallError %d %s\n", e.code, e.what());
        return 2;
    }
    return 0;
}
