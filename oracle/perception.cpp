// perception.cpp -- perception neighbourhood order for the CPU oracle (checker).
// TEST INFRASTRUCTURE ONLY.
//
// computeNeighboringSitesCentral (Percolation_FD/multiSPDP.cpp:1442-1461, same
// function at Rietkerk_FPE/rietkerk_bjork_FKE.cpp) lists the offsets (x, y) of
// [-R, R]^2 with x^2 + y^2 <= R^2, x outer and y inner, and sorts them with
// std::sort under a comparator of squared distance alone (comparePairs,
// multiSPDP.cpp:108-111).  std::sort is not stable: the order inside a shell of
// equal distance is what libstdc++'s introsort makes of that input sequence, and
// the reference's sums over the first n offsets depend on it.  The same call on
// the same sequence reproduces it; tests pin the result against the reference's
// own Test/test_perception_radii.cpp built from source.
#include <algorithm>
#include <cstdint>
#include <utility>
#include <vector>

extern "C" int tmo_perception_offsets(int R, int32_t* out, int cap)
{
    std::vector<std::pair<int, int>> v;
    for (int x = -R; x <= R; ++x)
        for (int y = -R; y <= R; ++y)
            if (x * x + y * y <= R * R) v.emplace_back(x, y);
    std::sort(v.begin(), v.end(), [](const std::pair<int, int>& a, const std::pair<int, int>& b) {
        return (a.first * a.first + a.second * a.second) < (b.first * b.first + b.second * b.second);
    });
    if ((int)v.size() > cap) return -1;
    for (size_t k = 0; k < v.size(); ++k) {
        out[2 * k] = v[k].first;
        out[2 * k + 1] = v[k].second;
    }
    return (int)v.size();
}
