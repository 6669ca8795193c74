// Uses the C++ mirror (include/sparkx_hb.hpp) the way a reference user would use
// HyperBallInner: build, read accumulators, compute and fetch centralities.
// Prints one line per result so the test can compare them with the Python mirror.
#include <cstdio>
#include <vector>

#include "../../include/sparkx_hb.hpp"

int main(int argc, char **argv)
{
    // 0-1-2-3-4 path plus a triangle 5-6-7, symmetric, u64 CSR like the reference's files
    std::vector<uint64_t> off = {0, 1, 3, 5, 7, 8, 10, 12, 14};
    std::vector<uint64_t> nbr = {1, 0, 2, 1, 3, 2, 4, 3, 6, 7, 5, 7, 5, 6};
    sparkx::CsrView g{off.data(), 8, 8, nbr.data(), 8, nbr.size()};
    try {
        auto hb = sparkx::HyperBall6::create(g);
        const auto dist = hb.geodesic_distances();
        const auto &c = hb.compute_closeness_centrality(true);
        const auto &lin = hb.get_or_compute_lins_centrality();
        printf("iters %zu\n", hb.get_iters());
        for (size_t u = 0; u < dist.size(); ++u) printf("node %zu %.17g %.17g %.17g\n", u, dist[u], c[u], lin[u]);
        try {
            (void)hb.get_hamonic_centrality(false);
            printf("error: cached harmonic should not exist\n");
            return 1;
        } catch (const sparkx::HyperBallError &) {
        }
        try {
            (void)hb.get_geodesic_distance(8);
            printf("error: out of range accepted\n");
            return 1;
        } catch (const std::out_of_range &) {
        }
    } catch (const sparkx::HyperBallError &e) {
        printf("HyperBallError %d %s\n", e.code, e.what());
        return 2;
    }
    (void)argc;
    (void)argv;
    return 0;
}
