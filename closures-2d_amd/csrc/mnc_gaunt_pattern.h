// mnc_gaunt_pattern.h -- compile-time sparsity of the Gaunt map for moment
// order N (host and device).
//
// The table's basis (build_tables) is the real spherical harmonics that are
// even in mu, ordered l = 0, 1, ...; m = -l..l with l + m even: cos(m phi) P_l^m
// for m > 0, sin(|m| phi) P_l^|m| for m < 0.  The product of basis functions
// (l1, m1) and (l2, m2) is a combination of order-2N basis functions (L, M)
// with |l1 - l2| <= L <= l1 + l2, L = l1 + l2 (mod 2), L >= |M|, and
// |M| in {|m1| + |m2|, ||m1| - |m2||} of the type the product-to-sum formulas
// give (cos cos and sin sin -> cos, sin cos -> sin, sin(0) dropped).  Those
// rules give a superset of the nonzero coefficients of the fitted map
// (momopt_init.cpp:224-259, restated in build_tables); at N = 4 and Q >= 12 it
// has 434 terms, two of them zero.  Kernels that walk this pattern with
// unrolled loops have every index at compile time; the coefficients are read
// from an array in pattern order (gaunt_pattern_values), 0 where the fit has
// no entry.
#pragma once

namespace mnc {

constexpr int gp_nm(int n) { return (n + 1) * (n + 2) / 2; }
constexpr int gp_iabs(int x) { return x < 0 ? -x : x; }

// flat index of (l, m) in the basis above
constexpr int gp_index(int l, int m) {
    int k = 0;
    for (int ll = 0; ll < l; ++ll) k += ll + 1;
    return k + (m + l) / 2;
}

constexpr void gp_lm(int k, int &l, int &m) {
    l = 0;
    while (k > l) { k -= l + 1; ++l; }
    m = 2 * k - l;
}

// Visits the pattern's terms in the order (i, j >= i, k ascending); the
// callback gets (i, j, k).  Returns the number of terms.
template <class F>
constexpr int gp_walk(int n, F &&f) {
    const int nm = gp_nm(n);
    int count = 0;
    for (int i = 0; i < nm; ++i) {
        for (int j = i; j < nm; ++j) {
            int l1 = 0, m1 = 0, l2 = 0, m2 = 0;
            gp_lm(i, l1, m1);
            gp_lm(j, l2, m2);
            const int a = gp_iabs(m1), b = gp_iabs(m2);
            const bool cosine = (m1 >= 0) == (m2 >= 0);
            bool hit[(2 * 16 + 1) * (2 * 16 + 2) / 2] = {};
            const int ms[2] = {a + b, gp_iabs(a - b)};
            for (int s = 0; s < 2; ++s) {
                const int M = ms[s];
                if (!cosine && M == 0) continue;
                const int m = cosine ? M : -M;
                for (int L = gp_iabs(l1 - l2); L <= l1 + l2; ++L) {
                    if ((L + l1 + l2) % 2 != 0 || L < M) continue;
                    hit[gp_index(L, m)] = true;
                }
            }
            const int nmm = gp_nm(2 * n);
            for (int k = 0; k < nmm; ++k)
                if (hit[k]) { f(i, j, k); ++count; }
        }
    }
    return count;
}

constexpr int gp_count(int n) {
    return gp_walk(n, [](int, int, int) {});
}

template <int N>
struct GauntPattern {
    static constexpr int nm = gp_nm(N);
    static constexpr int count = gp_count(N);
    short i[count] = {}, j[count] = {}, k[count] = {};
};

template <int N>
constexpr GauntPattern<N> gaunt_pattern() {
    GauntPattern<N> p{};
    int e = 0;
    gp_walk(N, [&](int i, int j, int k) {
        p.i[e] = short(i);
        p.j[e] = short(j);
        p.k[e] = short(k);
        ++e;
    });
    return p;
}

}  // namespace mnc
