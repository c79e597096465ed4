// wd_gen.cpp -- deterministic synthetic gauge and spinor fields (host code).
//
// Every value depends only on (seed, stream, global lexicographic site,
// component), so fields are identical for any thread count or volume split.
#include "wilson_gen.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <thread>
#include <vector>

namespace {

typedef std::complex<double> cd;

constexpr uint32_t kPhiloxM0 = 0xD2511F53u, kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u, kPhiloxW1 = 0xBB67AE85u;

inline void philox(const uint32_t c_in[4], const uint32_t k_in[2], uint32_t out[4])
{
    uint32_t c0 = c_in[0], c1 = c_in[1], c2 = c_in[2], c3 = c_in[3];
    uint32_t k0 = k_in[0], k1 = k_in[1];
    for (int round = 0; round < 10; ++round) {
        if (round > 0) {
            k0 += kPhiloxW0;
            k1 += kPhiloxW1;
        }
        const uint64_t p0 = (uint64_t)kPhiloxM0 * c0;
        const uint64_t p1 = (uint64_t)kPhiloxM1 * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

// 53-bit uniform in [0, 1) from two 32-bit words
inline double u01(uint32_t a, uint32_t b)
{
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

// Stream of Gaussian pairs for one (seed, stream, site).
struct SiteGauss {
    uint32_t key[2];
    uint32_t ctr[4];
    double pending;
    bool have;
    SiteGauss(uint64_t seed, uint32_t stream, uint64_t site)
    {
        key[0] = (uint32_t)seed;
        key[1] = (uint32_t)(seed >> 32);
        ctr[0] = (uint32_t)site;
        ctr[1] = (uint32_t)(site >> 32);
        ctr[2] = 0;
        ctr[3] = stream;
        have = false;
        pending = 0.0;
    }
    double next()
    {
        if (have) {
            have = false;
            return pending;
        }
        uint32_t r[4];
        philox(ctr, key, r);
        ++ctr[2];
        const double u1 = 1.0 - u01(r[0], r[1]); // (0, 1]
        const double u2 = u01(r[2], r[3]);
        const double rad = std::sqrt(-2.0 * std::log(u1));
        const double th = 6.283185307179586476925286766559 * u2;
        pending = rad * std::sin(th);
        have = true;
        return rad * std::cos(th);
    }
    cd cnext()
    {
        const double re = next();
        const double im = next();
        return cd(re, im);
    }
};

inline long lex_of(const int x[4], const int L[4])
{
    return x[0] + (long)L[0] * (x[1] + (long)L[1] * (x[2] + (long)L[2] * x[3]));
}

inline long eo_of(const int x[4], const int L[4])
{
    const long V = (long)L[0] * L[1] * L[2] * L[3];
    return lex_of(x, L) / 2 + (long)((x[0] + x[1] + x[2] + x[3]) & 1) * (V / 2);
}

inline void coords(long lex, const int L[4], int x[4])
{
    x[0] = (int)(lex % L[0]);
    lex /= L[0];
    x[1] = (int)(lex % L[1]);
    lex /= L[1];
    x[2] = (int)(lex % L[2]);
    x[3] = (int)(lex / L[2]);
}

template <class F>
void parallel_sites(long V, F fn)
{
    unsigned nt = std::thread::hardware_concurrency();
    nt = std::max(1u, std::min(nt, 64u));
    if (V < 4096)
        nt = 1;
    std::vector<std::thread> th;
    const long chunk = (V + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
        const long lo = t * chunk, hi = std::min(V, lo + chunk);
        if (lo >= hi)
            break;
        th.emplace_back([=] {
            for (long s = lo; s < hi; ++s)
                fn(s);
        });
    }
    for (auto &t : th)
        t.join();
}

// Gram-Schmidt of rows 0 and 1, row 2 = conj(row0 x row1): an SU(3) matrix.
void reunitarise(cd m[3][3])
{
    double n0 = 0.0;
    for (int c = 0; c < 3; ++c)
        n0 += std::norm(m[0][c]);
    n0 = 1.0 / std::sqrt(n0);
    for (int c = 0; c < 3; ++c)
        m[0][c] *= n0;
    cd ov = 0.0;
    for (int c = 0; c < 3; ++c)
        ov += std::conj(m[0][c]) * m[1][c];
    for (int c = 0; c < 3; ++c)
        m[1][c] -= ov * m[0][c];
    double n1 = 0.0;
    for (int c = 0; c < 3; ++c)
        n1 += std::norm(m[1][c]);
    n1 = 1.0 / std::sqrt(n1);
    for (int c = 0; c < 3; ++c)
        m[1][c] *= n1;
    for (int c = 0; c < 3; ++c) {
        const int a = (c + 1) % 3, b = (c + 2) % 3;
        m[2][c] = std::conj(m[0][a] * m[1][b] - m[0][b] * m[1][a]);
    }
}

void store_link(double *A, long eo, const cd m[3][3])
{
    double *d = A + 18 * eo;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            d[2 * (a * 3 + b)] = m[a][b].real();
            d[2 * (a * 3 + b) + 1] = m[a][b].imag();
        }
}

// exp(X) for a 3x3 complex X by scaling and squaring with a Taylor series
void expm3(const cd X[3][3], cd out[3][3])
{
    double nrm = 0.0;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            nrm = std::max(nrm, std::abs(X[a][b]));
    int s = 0;
    while (nrm * 3.0 > 0.5) {
        nrm *= 0.5;
        ++s;
    }
    const double sc = std::ldexp(1.0, -s);
    cd Y[3][3], term[3][3], acc[3][3];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            Y[a][b] = X[a][b] * sc;
            term[a][b] = (a == b) ? 1.0 : 0.0;
            acc[a][b] = term[a][b];
        }
    for (int k = 1; k <= 16; ++k) {
        cd nt[3][3];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                cd v = 0.0;
                for (int c = 0; c < 3; ++c)
                    v += term[a][c] * Y[c][b];
                nt[a][b] = v / (double)k;
            }
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                term[a][b] = nt[a][b];
                acc[a][b] += nt[a][b];
            }
    }
    for (int i = 0; i < s; ++i) {
        cd sq[3][3];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                cd v = 0.0;
                for (int c = 0; c < 3; ++c)
                    v += acc[a][c] * acc[c][b];
                sq[a][b] = v;
            }
        std::memcpy(acc, sq, sizeof(sq));
    }
    std::memcpy(out, acc, sizeof(acc));
}

} // namespace

extern "C" void wdg_philox4x32(const uint32_t c[4], const uint32_t k[2], uint32_t out[4])
{
    philox(c, k, out);
}

extern "C" void wdg_gauge_random_su3(const int L[4], uint64_t seed, double *const A[4])
{
    const long V = (long)L[0] * L[1] * L[2] * L[3];
    parallel_sites(V, [&](long lex) {
        int x[4];
        coords(lex, L, x);
        const long eo = eo_of(x, L);
        for (int mu = 0; mu < 4; ++mu) {
            SiteGauss g(seed, (uint32_t)mu, (uint64_t)lex);
            cd m[3][3];
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b)
                    m[a][b] = g.cnext();
            reunitarise(m);
            store_link(A[mu], eo, m);
        }
    });
}

extern "C" void wdg_gauge_near_identity(const int L[4], uint64_t seed, double eps,
                                        double *const A[4])
{
    const long V = (long)L[0] * L[1] * L[2] * L[3];
    parallel_sites(V, [&](long lex) {
        int x[4];
        coords(lex, L, x);
        const long eo = eo_of(x, L);
        for (int mu = 0; mu < 4; ++mu) {
            SiteGauss g(seed, (uint32_t)mu, (uint64_t)lex);
            cd a[3][3], H[3][3], X[3][3], U[3][3];
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c)
                    a[r][c] = g.cnext();
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c)
                    H[r][c] = 0.5 * (a[r][c] + std::conj(a[c][r]));
            const cd tr = (H[0][0] + H[1][1] + H[2][2]) / 3.0;
            for (int r = 0; r < 3; ++r)
                H[r][r] -= tr;
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c)
                    X[r][c] = cd(0.0, eps) * H[r][c];
            expm3(X, U);
            reunitarise(U);
            store_link(A[mu], eo, U);
        }
    });
}

extern "C" void wdg_gauge_unit(const int L[4], double *const A[4])
{
    const long V = (long)L[0] * L[1] * L[2] * L[3];
    for (int mu = 0; mu < 4; ++mu) {
        std::memset(A[mu], 0, sizeof(double) * 18 * V);
        for (long e = 0; e < V; ++e) {
            A[mu][18 * e + 0] = 1.0;
            A[mu][18 * e + 8] = 1.0;
            A[mu][18 * e + 16] = 1.0;
        }
    }
}

extern "C" void wdg_gauge_flip_boundary(const int L[4], int mu, double *const A[4])
{
    const long V = (long)L[0] * L[1] * L[2] * L[3];
    for (long lex = 0; lex < V; ++lex) {
        int x[4];
        coords(lex, L, x);
        if (x[mu] != L[mu] - 1)
            continue;
        double *d = A[mu] + 18 * eo_of(x, L);
        for (int k = 0; k < 18; ++k)
            d[k] = -d[k];
    }
}

extern "C" void wdg_spinor_gaussian(const int L[4], uint64_t seed, uint32_t stream, double *A)
{
    const long V = (long)L[0] * L[1] * L[2] * L[3];
    parallel_sites(V, [&](long lex) {
        int x[4];
        coords(lex, L, x);
        double *d = A + 24 * eo_of(x, L);
        SiteGauss g(seed, 0x10000u + stream, (uint64_t)lex);
        for (int k = 0; k < 24; ++k)
            d[k] = g.next();
    });
}

extern "C" void wdg_spinor_point(const int L[4], const int x[4], int spin, int colour, double *A)
{
    const long V = (long)L[0] * L[1] * L[2] * L[3];
    std::memset(A, 0, sizeof(double) * 24 * V);
    A[24 * eo_of(x, L) + 2 * (spin * 3 + colour)] = 1.0;
}
