// Work decomposition of the pair-symmetric update pass (k_sym_update) on the
// Hermitian half layout.  Shared by the kernel (device), the unit table
// builder (host, sym_units.cpp) and the host-side check of both
// (tests/sym_units_check.cpp), so that the check runs the kernel's own index
// arithmetic.  No HIP dependency when compiled as plain C++.
#pragma once

#include <cstdint>
#include <vector>

#if !defined(__HIPCC__) && !defined(__host__)
#define __host__
#define __device__
#endif

namespace g12 {

// Hermitian half layout of the pair-symmetric path: a G2 tensor keeps only its
// tiles (p1, q1) with p1 <= q1 (the tile (q1, p1) is the conjugate transpose),
// each an N x N block of rows p2 and columns q2, contiguous, block (p1, q1) at
// hblock(p1, q1) N^2.  Buffers of complex64 elements use the same element
// offsets.  Diagonal tiles are stored whole and kept exactly Hermitian.
__host__ __device__ inline uint32_t hblock(int p1, int q1) { return (uint32_t)(q1 * (q1 + 1) / 2 + p1); }

// The exchanged element X[p2][p1][q2][q1] of a stored element (p2, q2) of
// block (p1, q1) is element (p1, q1) of block (p2, q2) when p2 <= q2
// (direct), else -conj of element (q1, p1) of block (q2, p2) (mirrored).  The
// stored elements fall into two families, each closed under that pairing:
//   direct (type 0): elements (t, w), t <= w, of any block;
//   mirrored (type 1): elements (t, w), t > w, of the blocks p1 < q1.
// The lower elements of the diagonal blocks are the Hermitian images of their
// upper ones and are written with them.  A unit takes a band of rows of one
// side and the matching band of columns of the other, like a blocked
// transpose, so that every global access runs along rows:
//   direct (x = a <= y = m, band k): side 1 = row m, columns m..N-1 of the
//     blocks (a, q1), q1 in [max(a, 16k), min(16k + 16, N)); side 2 = row a,
//     columns q1 of the blocks (m, q2), q2 = m..N-1;
//   mirrored (x = b <= y = n, band k): side 1 = row n, columns 0..n-1 of the
//     blocks (p1, b), p1 in [16k, min(16k + 16, b)); side 2 = row b, columns
//     p1 of the blocks (p2, n), p2 = 0..n-1.
// Element (i, j) of side 1 pairs with element (j, i) of side 2; with x = y
// side 2 is the side-1 family itself and is read for X only.
constexpr int SU_THREADS = 1024;  // slots per block
constexpr int SU_MAXU = 32;       // units per block
constexpr int SU_LA = 1536;       // LDS elements per block, side 1 (R x (L + 1) per unit)
constexpr int SU_LB = 2560;       // side 2 (L x (R + 1) per unit)

// A unit: type, x, y, band k, slot offset so in its block and LDS offsets la,
// lb of its two sides; R rows of L elements on side 1 from row r0 of the band.
struct SuUnit {
    int type, x, y, k, so, la, lb;
    int r0, R, L;
};
__host__ __device__ inline uint64_t su_encode(int type, int x, int y, int k, int so, int la, int lb) {
    return (uint64_t)so | ((uint64_t)la << 11) | ((uint64_t)lb << 22) | ((uint64_t)type << 34) |
           ((uint64_t)x << 35) | ((uint64_t)y << 42) | ((uint64_t)k << 49);
}
__host__ __device__ inline int su_slot(uint64_t w) { return (int)(w & 2047u); }
__host__ __device__ inline SuUnit su_decode(uint64_t w, int N) {
    SuUnit u;
    u.so = su_slot(w);
    u.la = (int)((w >> 11) & 2047u);
    u.lb = (int)((w >> 22) & 4095u);
    u.type = (int)((w >> 34) & 1u);
    u.x = (int)((w >> 35) & 127u);
    u.y = (int)((w >> 42) & 127u);
    u.k = (int)((w >> 49) & 7u);
    const int top = 16 * u.k + 16;
    if (u.type == 0) {
        u.r0 = u.x > 16 * u.k ? u.x : 16 * u.k;
        u.R = (top < N ? top : N) - u.r0;
        u.L = N - u.y;
    } else {
        u.r0 = 16 * u.k;
        u.R = (top < u.x ? top : u.x) - u.r0;
        u.L = u.y;
    }
    return u;
}

// Index of the unit that holds slot t of a block whose cnt units (increasing
// slot offsets, the first at 0) are ud[0 .. cnt): the last one starting at or
// before t.
__host__ __device__ inline int su_find(const uint64_t* ud, int cnt, int t) {
    int lo = 0, hi = cnt - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (su_slot(ud[mid]) <= t) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// The element pair of slot e (< R L) of a unit: side 1 is row i1 of R,
// column j1 of L; side 2 row j2 of L, column i2 of R.  o: element offsets in
// the half layout; (bp, bq): the block; z >= 0: the element is the diagonal
// element (z, z) of its block (it carries D[bp][bq][z] and D[bq][bp][z]); im:
// the offset of its Hermitian image in a diagonal block, or SU_NONE; w2:
// side 2 is written (x != y).
constexpr uint32_t SU_NONE = 0xffffffffu;
struct SuElem {
    int i1, j1, i2, j2;
    uint32_t o1, o2, im1, im2;
    int bp1, bq1, bp2, bq2, z1, z2;
    bool w2;
};
__host__ __device__ inline SuElem su_elem(const SuUnit& u, int e, int N) {
    SuElem r;
    r.i1 = e / u.L;
    r.j1 = e - r.i1 * u.L;
    r.j2 = e / u.R;
    r.i2 = e - r.j2 * u.R;
    const uint32_t N2 = (uint32_t)(N * N);
    r.im1 = r.im2 = SU_NONE;
    r.z1 = r.z2 = -1;
    r.w2 = u.x != u.y;
    if (u.type == 0) {
        const int a = u.x, m = u.y;
        r.bp1 = a;
        r.bq1 = u.r0 + r.i1;
        r.bp2 = m;
        r.bq2 = m + r.j2;
        r.o1 = hblock(r.bp1, r.bq1) * N2 + (uint32_t)(m * N + m + r.j1);
        r.o2 = hblock(r.bp2, r.bq2) * N2 + (uint32_t)(a * N + u.r0 + r.i2);
        if (r.j1 == 0) r.z1 = m;
        if (u.r0 + r.i2 == a) r.z2 = a;
        if (r.bp1 == r.bq1 && r.j1 > 0) r.im1 = hblock(a, a) * N2 + (uint32_t)((m + r.j1) * N + m);
        if (r.j2 == 0 && u.r0 + r.i2 != a) r.im2 = hblock(m, m) * N2 + (uint32_t)((u.r0 + r.i2) * N + a);
    } else {
        const int b = u.x, n = u.y;
        r.bp1 = u.r0 + r.i1;
        r.bq1 = b;
        r.bp2 = r.j2;
        r.bq2 = n;
        r.o1 = hblock(r.bp1, r.bq1) * N2 + (uint32_t)(n * N + r.j1);
        r.o2 = hblock(r.bp2, r.bq2) * N2 + (uint32_t)(b * N + u.r0 + r.i2);
    }
    return r;
}

// Unit table for size N: every unit, packed into blocks of SU_THREADS slots
// (best fit, largest units first, at most SU_MAXU units and SU_LA / SU_LB LDS
// elements per block); units[first[b] .. first[b + 1]) belong to block b, in
// increasing slot offset.  Returns the number of blocks (per spin).
int64_t su_table(int N, std::vector<uint64_t>& units, std::vector<uint32_t>& first);

}  // namespace g12
