// mnc_tables.cpp -- quadrature, real spherical-harmonic table and Gaunt map
// for the Mn closure (the host setup of MomOptSolver::init,
// src/momopt/momopt_init.cpp:88-259, and utils_getGaussianWeightsAndNodes,
// src/utils.cpp:55-60).  Host code; the tables are uploaded once per context.
#include <cmath>
#include <cstring>

#include "mnc_internal.h"
#include "mnc_gaunt_pattern.h"

namespace mnc {
namespace {

// Gauss-Legendre nodes/weights on [-1,1] in ascending order.  Each positive
// root of P_n is refined by Newton's method on the three-term recurrence.
void gauss_legendre(int n, std::vector<double> &x, std::vector<double> &w) {
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    auto legendre = [n](double z, double &pn, double &dpn) {
        double pm1 = 1.0, p = z;
        if (n == 0) { pn = 1.0; dpn = 0.0; return; }
        for (int k = 2; k <= n; ++k) {
            double pk = ((2.0 * k - 1.0) * z * p - (k - 1.0) * pm1) / k;
            pm1 = p;
            p = pk;
        }
        pn = p;
        dpn = n * (z * p - pm1) / (z * z - 1.0);
    };
    for (int i = 0; i < (n + 1) / 2; ++i) {
        double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
        double pn = 0.0, dpn = 1.0;
        for (int it = 0; it < 100; ++it) {
            legendre(z, pn, dpn);
            double dz = pn / dpn;
            z -= dz;
            if (std::fabs(dz) < 1e-16 * std::fabs(z) && it >= 3) break;
        }
        legendre(z, pn, dpn);
        double wt = 2.0 / ((1.0 - z * z) * dpn * dpn);
        x[i] = -z;
        x[n - 1 - i] = z;
        w[i] = w[n - 1 - i] = wt;
    }
    if (n % 2) x[n / 2] = 0.0;
}

// Fully normalised associated Legendre function including the Condon-Shortley
// phase: sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!) P_l^m(x), by upward recursion in l
// from the sectoral value.
double sph_plm(int l, int m, double x) {
    double s = std::sqrt((1.0 - x) * (1.0 + x));
    // Sectoral value (-1)^m sqrt((2m+1)/4pi) sqrt((2m-1)!!/(2m)!!) s^m, built as
    // 1/sqrt(4pi) * prod_{k<=m} (-s) sqrt((2k+1)/(2k)), which telescopes to it.
    double pmm = 1.0 / std::sqrt(4.0 * M_PI);
    for (int k = 1; k <= m; ++k) pmm *= -s * std::sqrt((2.0 * k + 1.0) / (2.0 * k));
    if (l == m) return pmm;
    double p1 = x * std::sqrt(2.0 * m + 3.0) * pmm;
    if (l == m + 1) return p1;
    double p0 = pmm;
    for (int ll = m + 2; ll <= l; ++ll) {
        double l2 = double(ll) * ll, m2 = double(m) * m;
        double a = std::sqrt((4.0 * l2 - 1.0) / (l2 - m2));
        double b = std::sqrt(((ll - 1.0) * (ll - 1.0) - m2) / (4.0 * (ll - 1.0) * (ll - 1.0) - 1.0));
        double pl = a * (x * p1 - b * p0);
        p0 = p1;
        p1 = pl;
    }
    return p1;
}

}  // namespace

int build_tables(int moment_order, int quad_order, int use_gaunt, Tables &t) {
    if (moment_order < 0 || moment_order > 16 || quad_order < 2 || quad_order % 2 != 0)
        return MNC_ERR_ARG;
    const int N = moment_order, Q = quad_order;
    t.N = N;
    t.Q = Q;
    t.use_gaunt = use_gaunt ? 1 : 0;
    t.nq = Q * Q;                                    // momopt_init.cpp:88
    t.nm = (N + 1) * (N + 2) / 2;                    // :89
    t.nmm = t.use_gaunt ? (2 * N + 1) * (2 * N + 2) / 2 : t.nm;   // :90-93
    t.np2 = t.nmm;
    const int nq = t.nq, np2 = t.np2;
    t.w.assign(nq, 0.0);
    t.p.assign(size_t(nq) * np2, 0.0);
    t.xi.assign(nq, 0.0);
    t.eta.assign(nq, 0.0);
    std::vector<double> mu, wg;
    gauss_legendre(Q, mu, wg);
    const int nsh = t.use_gaunt ? 2 * N + 1 : N + 1;
    // Hemisphere mu < 0 (the first Q/2 ascending nodes), 2Q azimuths, weights
    // 2 pi w_i / Q: the integrands are even in mu, so this integrates over the
    // whole sphere (momopt_init.cpp:176-217).
    for (int q1 = 0; q1 < Q / 2; ++q1) {
        const double wq = 2.0 * M_PI / Q * wg[q1];
        const double m_u = mu[q1];
        const double st = std::sqrt(1.0 - m_u * m_u);
        for (int q2 = 0; q2 < 2 * Q; ++q2) {
            const double phi = (q2 + 0.5) * M_PI / Q;
            const int q = 2 * q1 * Q + q2;
            t.w[q] = wq;
            t.xi[q] = st * std::cos(phi);
            t.eta[q] = st * std::sin(phi);
            int k = 0;
            for (int l = 0; l < nsh; ++l) {
                for (int m = -l; m <= l; ++m) {
                    if ((m + l) % 2) continue;
                    double v;
                    if (m < 0) v = M_SQRT2 * std::sin(m * phi) * sph_plm(l, -m, m_u);
                    else if (m > 0) v = M_SQRT2 * std::cos(m * phi) * sph_plm(l, m, m_u);
                    else v = sph_plm(l, 0, m_u);
                    t.p[size_t(q) * np2 + k++] = v;
                }
            }
        }
    }
    t.gptr.clear();
    t.gk.clear();
    t.gv.clear();
    if (t.use_gaunt) {
        // Least-squares fit of m_i m_j on the nmm order-2N harmonics at the nq
        // nodes (momopt_init.cpp:221-259), by modified Gram-Schmidt QR.
        const int nm = t.nm, nmm = t.nmm;
        if (nq < nmm) return MNC_ERR_ARG;
        std::vector<double> Qm(size_t(nq) * nmm), R(size_t(nmm) * nmm, 0.0);
        for (int q = 0; q < nq; ++q)
            for (int k = 0; k < nmm; ++k) Qm[size_t(k) * nq + q] = t.p[size_t(q) * np2 + k];
        for (int k = 0; k < nmm; ++k) {
            double *qk = &Qm[size_t(k) * nq];
            for (int j = 0; j < k; ++j) {
                const double *qj = &Qm[size_t(j) * nq];
                double s = 0.0;
                for (int q = 0; q < nq; ++q) s += qj[q] * qk[q];
                R[size_t(j) * nmm + k] = s;
                for (int q = 0; q < nq; ++q) qk[q] -= s * qj[q];
            }
            double nrm = 0.0;
            for (int q = 0; q < nq; ++q) nrm += qk[q] * qk[q];
            nrm = std::sqrt(nrm);
            if (nrm == 0.0) return MNC_ERR_ARG;
            R[size_t(k) * nmm + k] = nrm;
            for (int q = 0; q < nq; ++q) qk[q] /= nrm;
        }
        std::vector<std::vector<int>> ks(size_t(nm) * nm);
        std::vector<std::vector<double>> vs(size_t(nm) * nm);
        std::vector<double> b(nq), c(nmm), beta(nmm);
        for (int i = 0; i < nm; ++i) {
            for (int j = i; j < nm; ++j) {
                for (int q = 0; q < nq; ++q)
                    b[q] = t.p[size_t(q) * np2 + i] * t.p[size_t(q) * np2 + j];
                for (int k = 0; k < nmm; ++k) {   // c = Q^T b (MGS order)
                    const double *qk = &Qm[size_t(k) * nq];
                    double s = 0.0;
                    for (int q = 0; q < nq; ++q) s += qk[q] * b[q];
                    c[k] = s;
                    for (int q = 0; q < nq; ++q) b[q] -= s * qk[q];
                }
                for (int k = nmm - 1; k >= 0; --k) {
                    double s = c[k];
                    for (int jj = k + 1; jj < nmm; ++jj) s -= R[size_t(k) * nmm + jj] * beta[jj];
                    beta[k] = s / R[size_t(k) * nmm + k];
                }
                for (int k = 0; k < nmm; ++k) {
                    if (std::fabs(beta[k]) > 10e-10) {   // momopt_init.cpp:249
                        ks[size_t(i) * nm + j].push_back(k);
                        vs[size_t(i) * nm + j].push_back(beta[k]);
                    }
                }
                if (j != i) {
                    ks[size_t(j) * nm + i] = ks[size_t(i) * nm + j];
                    vs[size_t(j) * nm + i] = vs[size_t(i) * nm + j];
                }
            }
        }
        t.gptr.resize(size_t(nm) * nm + 1);
        int off = 0;
        for (int pidx = 0; pidx < nm * nm; ++pidx) {
            t.gptr[pidx] = off;
            for (size_t c2 = 0; c2 < ks[pidx].size(); ++c2) {
                t.gk.push_back(ks[pidx][c2]);
                t.gv.push_back(vs[pidx][c2]);
                ++off;
            }
        }
        t.gptr[size_t(nm) * nm] = off;
    }
    return MNC_OK;
}

namespace {

// Round-to-nearest-even bf16 of a float (finite inputs).
unsigned short bf16_rne(float f) {
    unsigned u;
    std::memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
float bf16_val(unsigned short b) {
    const unsigned u = unsigned(b) << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

}  // namespace

// Operand tables of the mixed-precision sweep.  Quadrature point q of a 16-point
// tile T sits at row q - 16T of the MFMA operands; the lane layouts are those of
// v_mfma_f64_16x16x4f64 (lane l: A[l%16][l/16], B[l/16][l%16], result element r
// at row R(r, l/16) = mfma_f64_row) and v_mfma_f32_16x16x32_bf16 (lane l:
// A[l%16][8(l/16)+t], B[8(l/16)+t][l%16]).  The exponent z = alpha.m comes out
// of the first product with element r of lane l at point 16T + R(r, l/16); the
// later products take their K index in that order:
//   z[T][kk][l] = p[16T + l%16][4kk + l/16]                  (alpha.m, A operand)
//   g[T][r][l]  = w_q p[q][l%16], q = 16T + R(r, l/16)       (sums, B operand;
//                 column 15 holds w_q: the sum of w e)
//   h[kb][nt][hl][l][t] = hi/lo bf16 of w_q p2N[q][16nt + l%16],
//                 q = 16(2kb + t/4) + R(t%4, l/16)            (Hessian moments, B)
int build_mix_tables(int moment_order, int quad_order, MixTables &m) {
    Tables t;
    int st = build_tables(moment_order, quad_order, 1, t);   // order-2N table + Gaunt map
    if (st != MNC_OK) return st;
    const int nm = t.nm, nq = t.nq, np2 = t.np2;
    if (nm > 15) return MNC_ERR_ARG;
    const int nqp = (nq + 31) / 32 * 32;
    if (nqp / 16 > kMixMaxTiles) return MNC_ERR_ARG;
    m.nt = nqp / 16;
    m.nmm = t.nmm;
    m.nnt = (t.nmm + 15) / 16;
    auto P = [&](int q, int k) { return (q < nq && k < np2) ? t.p[size_t(q) * np2 + k] : 0.0; };
    auto W = [&](int q) { return q < nq ? t.w[q] : 0.0; };
    m.z.assign(size_t(m.nt) * 4 * 64, 0.0);
    m.g.assign(size_t(m.nt) * 4 * 64, 0.0);
    for (int T = 0; T < m.nt; ++T)
        for (int k4 = 0; k4 < 4; ++k4)
            for (int l = 0; l < 64; ++l) {
                const int kz = 4 * k4 + l / 16;
                m.z[(size_t(T) * 4 + k4) * 64 + l] = kz < nm ? P(16 * T + l % 16, kz) : 0.0;
                const int q = 16 * T + mfma_f64_row(k4, l / 16), j = l % 16;
                m.g[(size_t(T) * 4 + k4) * 64 + l] = j < nm ? W(q) * P(q, j) : (j == 15 ? W(q) : 0.0);
            }
    // three bf16 parts per table value: hi, then the rounded remainders mid
    // and lo (the two-part kernels read hi and mid)
    const int nkb = m.nt / 2;
    m.h.assign(size_t(nkb) * m.nnt * kMixParts * 64 * 8, 0);
    for (int kb = 0; kb < nkb; ++kb)
        for (int nt = 0; nt < m.nnt; ++nt)
            for (int l = 0; l < 64; ++l)
                for (int tt = 0; tt < 8; ++tt) {
                    const int q = 16 * (2 * kb + tt / 4) + mfma_f64_row(tt % 4, l / 16), col = 16 * nt + l % 16;
                    const float v = (float)(col < t.nmm ? W(q) * P(q, col) : 0.0);
                    const unsigned short hi = bf16_rne(v);
                    const float r1 = v - bf16_val(hi);
                    const unsigned short mid = bf16_rne(r1);
                    const unsigned short lo = bf16_rne(r1 - bf16_val(mid));
                    const size_t base = ((size_t(kb) * m.nnt + nt) * kMixParts) * 64 * 8;
                    m.h[base + size_t(l) * 8 + tt] = hi;
                    m.h[base + 64 * 8 + size_t(l) * 8 + tt] = mid;
                    m.h[base + 2 * 64 * 8 + size_t(l) * 8 + tt] = lo;
                }
    m.gptr = t.gptr;
    m.gk = t.gk;
    m.gv = t.gv;
    if (!gaunt_pattern_values(moment_order, t.gptr, t.gk, t.gv, m.pv)) m.pv.clear();
    return MNC_OK;
}

bool gaunt_pattern_values(int n, const std::vector<int> &gptr, const std::vector<int> &gk,
                          const std::vector<double> &gv, std::vector<float> &out) {
    const int nm = gp_nm(n);
    out.clear();
    if (gptr.size() != size_t(nm) * nm + 1) return false;
    size_t matched = 0, upper = 0;
    for (int i = 0; i < nm; ++i)
        for (int j = i; j < nm; ++j) upper += size_t(gptr[size_t(i) * nm + j + 1] - gptr[size_t(i) * nm + j]);
    gp_walk(n, [&](int i, int j, int k) {
        float v = 0.0f;
        const int pr = i * nm + j;
        for (int cc = gptr[pr]; cc < gptr[pr + 1]; ++cc)
            if (gk[cc] == k) { v = float(gv[cc]); ++matched; }
        out.push_back(v);
    });
    if (matched != upper) { out.clear(); return false; }
    return true;
}

// Operand tables of the FP64 matrix-core sweep (v_mfma_f64_16x16x4f64 layouts
// as for the mixed sweep; the exponent of point 16T + R(r, l/16) sits in
// element r of lane l after the first product):
//   z[T][kk][l]     = p[16T + l%16][4kk + l/16]                         (A operand)
//   g[T][r][nt][l]  = w_q p[q][16nt + l%16], column 16 ngt - 1 = w_q      (B operand)
//   h[T][r][nt][l]  = w_q p2N[q][16nt + l%16]                            (B operand)
// with q = 16T + R(r, l/16).
int build_g64_tables(int moment_order, int quad_order, G64Tables &m) {
    Tables t;
    int st = build_tables(moment_order, quad_order, 1, t);
    if (st != MNC_OK) return st;
    const int nm = t.nm, nq = t.nq, np2 = t.np2;
    m.nt = (nq + 15) / 16;
    m.kz = (nm + 3) / 4;
    m.ngt = (nm + 1 + 15) / 16;
    m.nht = (t.nmm + 15) / 16;
    m.nmm = t.nmm;
    auto P = [&](int q, int k) { return (q < nq && k < np2) ? t.p[size_t(q) * np2 + k] : 0.0; };
    auto W = [&](int q) { return q < nq ? t.w[q] : 0.0; };
    m.z.assign(size_t(m.nt) * m.kz * 64, 0.0);
    m.g.assign(size_t(m.nt) * 4 * m.ngt * 64, 0.0);
    m.h.assign(size_t(m.nt) * 4 * m.nht * 64, 0.0);
    for (int T = 0; T < m.nt; ++T) {
        for (int kk = 0; kk < m.kz; ++kk)
            for (int l = 0; l < 64; ++l) {
                const int k = 4 * kk + l / 16;
                m.z[(size_t(T) * m.kz + kk) * 64 + l] = k < nm ? P(16 * T + l % 16, k) : 0.0;
            }
        for (int r = 0; r < 4; ++r)
            for (int l = 0; l < 64; ++l) {
                const int q = 16 * T + mfma_f64_row(r, l / 16);
                for (int nt = 0; nt < m.ngt; ++nt) {
                    const int j = 16 * nt + l % 16;
                    m.g[((size_t(T) * 4 + r) * m.ngt + nt) * 64 + l] =
                        j < nm ? W(q) * P(q, j) : (j == 16 * m.ngt - 1 ? W(q) : 0.0);
                }
                for (int nt = 0; nt < m.nht; ++nt) {
                    const int col = 16 * nt + l % 16;
                    m.h[((size_t(T) * 4 + r) * m.nht + nt) * 64 + l] = col < t.nmm ? W(q) * P(q, col) : 0.0;
                }
            }
    }
    m.gptr = t.gptr;
    m.gk = t.gk;
    m.gv = t.gv;
    return MNC_OK;
}

}  // namespace mnc
