// Host-side check of the pair-symmetric update's work decomposition
// (g12kba_amd/csrc/sym_units.h, sym_units.cpp): the unit tables of N = 1..64
// and, through su_find / su_decode / su_elem, the kernel's own per-thread
// index arithmetic.  For every N: every block keeps to its slot, unit and LDS
// limits; every stored element of the Hermitian half layout, and every
// partial-trace entry D[X][Y][Z], is written exactly once; every access stays
// inside its buffer and its unit's LDS regions.  For N <= 12: the pairs give
// C = X + P(X) exactly, for a random anti-Hermitian X held in the half layout.
// Prints "ok" and exits 0, or names the first failure and exits 1.
#include <complex>
#include <cstdio>
#include <random>
#include <vector>

#include "sym_units.h"

using namespace g12;
typedef std::complex<double> cd;

static int fail(const char* what, int N, long long at) {
    std::printf("FAIL %s N=%d at %lld\n", what, N, at);
    return 1;
}

int main() {
    for (int N = 1; N <= 64; ++N) {
        std::vector<uint64_t> units;
        std::vector<uint32_t> first;
        const long long nb = su_table(N, units, first);
        const uint32_t N2 = (uint32_t)(N * N), tot = hblock(0, N) * N2;
        std::vector<int> cnt(tot, 0), dcnt((size_t)N * N * N, 0);
        // values: X in the half layout, anti-Hermitian diagonal blocks
        const bool vals = N <= 12;
        std::vector<cd> Xh(vals ? tot : 0), Cu(vals ? tot : 0);
        std::mt19937_64 rng(1234 + N);
        std::normal_distribution<double> nd;
        if (vals) {
            for (auto& v : Xh) v = cd(nd(rng), nd(rng));
            for (int p = 0; p < N; ++p)
                for (int t = 0; t < N; ++t)
                    for (int w = 0; w <= t; ++w) {
                        cd& up = Xh[hblock(p, p) * N2 + w * N + t];
                        if (w == t) up = cd(0.0, up.imag());
                        Xh[hblock(p, p) * N2 + t * N + w] = -std::conj(up);
                    }
        }
        auto full = [&](int p1, int p2, int q1, int q2) {  // X[p1][p2][q1][q2] from the half layout
            return p1 <= q1 ? Xh[hblock(p1, q1) * N2 + p2 * N + q2] : -std::conj(Xh[hblock(q1, p1) * N2 + q2 * N + p2]);
        };
        for (long long b = 0; b < nb; ++b) {
            const int c = (int)(first[b + 1] - first[b]);
            if (c < 1 || c > SU_MAXU) return fail("units per block", N, b);
            const uint64_t* ud = units.data() + first[b];
            if (su_slot(ud[0]) != 0) return fail("first slot", N, b);
            int used = 0, la = 0, lb = 0;
            for (int k = 0; k < c; ++k) {
                const SuUnit u = su_decode(ud[k], N);
                if (u.R < 1 || u.L < 1) return fail("empty unit", N, b);
                if (u.so != used || u.la != la || u.lb != lb) return fail("unit offsets", N, b);
                used += u.R * u.L;
                la += u.R * (u.L + 1);
                lb += u.L * (u.R + 1);
            }
            if (used > SU_THREADS || la > SU_LA || lb > SU_LB) return fail("block limits", N, b);
            // pass 1 as the kernel's gather: each live slot's two X elements into
            // its unit's LDS regions; pass 2: the update from the partners in LDS
            std::vector<cd> LA(SU_LA), LB(SU_LB);
            std::vector<char> fa(SU_LA, 0), fb(SU_LB, 0);
            for (int pass = 0; pass < 2; ++pass)
                for (int t = 0; t < SU_THREADS; ++t) {
                    const SuUnit u = su_decode(ud[su_find(ud, c, t)], N);
                    const int e = t - u.so;
                    if (e < 0) return fail("slot before its unit", N, b);
                    if (e >= u.R * u.L) continue;
                    const SuElem q = su_elem(u, e, N);
                    if (q.i1 >= u.R || q.j1 >= u.L || q.i2 >= u.R || q.j2 >= u.L) return fail("LDS index", N, b);
                    const int a1 = u.la + q.i1 * (u.L + 1) + q.j1, b2 = u.lb + q.j2 * (u.R + 1) + q.i2;
                    const int pa = u.lb + q.j1 * (u.R + 1) + q.i1, pb = u.la + q.i2 * (u.L + 1) + q.j2;
                    if (pass == 0) {
                        if (q.bp1 > q.bq1 || q.bp2 > q.bq2 || q.bq1 >= N || q.bq2 >= N || q.o1 >= tot || q.o2 >= tot)
                            return fail("element outside the half layout", N, b);
                        if (a1 >= SU_LA || b2 >= SU_LB || fa[a1] || fb[b2]) return fail("LDS slot", N, b);
                        fa[a1] = fb[b2] = 1;
                        if (vals) {
                            LA[a1] = Xh[q.o1];
                            LB[b2] = Xh[q.o2];
                        }
                        continue;
                    }
                    if (!fb[pa] || !fa[pb]) return fail("partner slot never written", N, b);
                    auto mark = [&](uint32_t o, uint32_t im, int bp, int bq, int z) {
                        ++cnt[o];
                        if (im != SU_NONE) {
                            if (im >= tot) return false;
                            ++cnt[im];
                        }
                        if (z >= 0) {
                            ++dcnt[((size_t)bp * N + bq) * N + z];
                            if (bp != bq) ++dcnt[((size_t)bq * N + bp) * N + z];
                        }
                        return true;
                    };
                    if (!mark(q.o1, q.im1, q.bp1, q.bq1, q.z1)) return fail("image", N, b);
                    if (q.w2 && !mark(q.o2, q.im2, q.bp2, q.bq2, q.z2)) return fail("image", N, b);
                    if (vals) {  // C = X + exchanged X (mirrored family: - conj), as the kernel forms it
                        auto sig = [&](cd v) { return u.type ? -std::conj(v) : v; };
                        Cu[q.o1] = LA[a1] + sig(LB[pa]);
                        if (q.im1 != SU_NONE) Cu[q.im1] = -std::conj(Cu[q.o1]);
                        if (q.w2) {
                            Cu[q.o2] = LB[b2] + sig(LA[pb]);
                            if (q.im2 != SU_NONE) Cu[q.im2] = -std::conj(Cu[q.o2]);
                        }
                    }
                }
        }
        for (uint32_t k = 0; k < tot; ++k)
            if (cnt[k] != 1) return fail("element not written exactly once", N, k);
        for (size_t k = 0; k < dcnt.size(); ++k)
            if (dcnt[k] != 1) return fail("D entry not written exactly once", N, (long long)k);
        if (vals)
            for (int p1 = 0; p1 < N; ++p1)
                for (int q1 = p1; q1 < N; ++q1)
                    for (int p2 = 0; p2 < N; ++p2)
                        for (int q2 = 0; q2 < N; ++q2) {
                            const cd ref = full(p1, p2, q1, q2) + full(p2, p1, q2, q1);
                            if (Cu[hblock(p1, q1) * N2 + p2 * N + q2] != ref) return fail("C value", N, p1 * N + q1);
                        }
    }
    std::printf("ok\n");
    return 0;
}
