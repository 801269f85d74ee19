// Unit table of the pair-symmetric update pass (k_sym_update): every unit of
// the two families of sym_units.h for one N, packed into blocks.
#include "sym_units.h"

#include <algorithm>

namespace g12 {

int64_t su_table(int N, std::vector<uint64_t>& units, std::vector<uint32_t>& first) {
    units.clear();
    first.clear();
    struct U {
        int type, x, y, k, R, L;
    };
    std::vector<U> us;
    const int nbd = (N + 15) / 16;
    for (int m = 0; m < N; ++m)  // direct: a <= m, bands from the one holding a
        for (int a = 0; a <= m; ++a)
            for (int k = a / 16; k < nbd; ++k) {
                const SuUnit v = su_decode(su_encode(0, a, m, k, 0, 0, 0), N);
                us.push_back({0, a, m, k, v.R, v.L});
            }
    for (int n = 1; n < N; ++n)  // mirrored: 1 <= b <= n, rows p1 < b
        for (int bx = 1; bx <= n; ++bx)
            for (int k = 0; k < (bx + 15) / 16; ++k) {
                const SuUnit v = su_decode(su_encode(1, bx, n, k, 0, 0, 0), N);
                us.push_back({1, bx, n, k, v.R, v.L});
            }
    // best fit, largest units first, under the slot, unit-count and LDS limits
    std::stable_sort(us.begin(), us.end(), [](const U& p, const U& q) { return p.R * p.L > q.R * q.L; });
    struct Bin {
        int used = 0, la = 0, lb = 0;
        std::vector<int> members;
    };
    std::vector<Bin> bins;
    std::vector<std::vector<int>> open(SU_THREADS + 1);  // bins by free slots
    for (int i = 0; i < (int)us.size(); ++i) {
        const U& v = us[i];
        const int sz = v.R * v.L, ea = v.R * (v.L + 1), eb = v.L * (v.R + 1);
        int bi = -1;
        for (int c = sz; c <= SU_THREADS && bi < 0; ++c)
            for (int j = (int)open[c].size() - 1; j >= 0; --j) {
                const Bin& q = bins[open[c][j]];
                if (q.la + ea <= SU_LA && q.lb + eb <= SU_LB) {
                    bi = open[c][j];
                    open[c].erase(open[c].begin() + j);
                    break;
                }
            }
        if (bi < 0) {
            bi = (int)bins.size();
            bins.emplace_back();
        }
        Bin& q = bins[bi];
        q.used += sz;
        q.la += ea;
        q.lb += eb;
        q.members.push_back(i);
        if ((int)q.members.size() < SU_MAXU) open[SU_THREADS - q.used].push_back(bi);
    }
    for (const Bin& q : bins) {
        first.push_back((uint32_t)units.size());
        int so = 0, oa = 0, ob = 0;
        for (int i : q.members) {
            const U& v = us[i];
            units.push_back(su_encode(v.type, v.x, v.y, v.k, so, oa, ob));
            so += v.R * v.L;
            oa += v.R * (v.L + 1);
            ob += v.L * (v.R + 1);
        }
    }
    first.push_back((uint32_t)units.size());
    return (int64_t)bins.size();
}

}  // namespace g12
