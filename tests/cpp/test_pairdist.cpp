// Host-side check of the Myers / packed-Hamming primitives used by the device
// kernels, against a textbook Wagner-Fischer DP and a byte loop.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#include "../../immunodbscan_amd/csrc/pairdist.h"

static uint64_t st = 12345;
static uint32_t rnd(uint32_t n)
{
    st = st * 6364136223846793005ull + 1442695040888963407ull;
    return (uint32_t)((st >> 33) % n);
}

static int wf(const uint8_t *a, int la, const uint8_t *b, int lb)
{
    std::vector<int> row(lb + 1);
    for (int j = 0; j <= lb; ++j) row[j] = j;
    for (int i = 1; i <= la; ++i) {
        int diag = row[0];
        row[0] = i;
        for (int j = 1; j <= lb; ++j) {
            int up = row[j], best = diag + (a[i - 1] != b[j - 1]);
            if (up + 1 < best) best = up + 1;
            if (row[j - 1] + 1 < best) best = row[j - 1] + 1;
            row[j] = best;
            diag = up;
        }
    }
    return row[lb];
}

template <int W>
static int myers(const uint8_t *pat, int m, const uint8_t *txt, int n, int K)
{
    std::vector<uint32_t> peq((size_t)K * W, 0);
    for (int i = 0; i < m; ++i) peq[(size_t)pat[i] * W + (i >> 5)] |= 1u << (i & 31);
    idb_myers<W> s;
    idb_myers_init<W>(s, (uint32_t)m);
    if (m == 0) return n;
    const uint32_t tw = (uint32_t)(m - 1) >> 5, tb = (uint32_t)(m - 1) & 31;
    for (int j = 0; j < n; ++j) idb_myers_step<W>(s, &peq[(size_t)txt[j] * W], tw, tb);
    return s.score;
}

template <int W>
static int check_w(int trials, int K)
{
    int bad = 0;
    uint8_t a[32 * W + 8], b[32 * W + 8];
    for (int t = 0; t < trials; ++t) {
        int m = (int)rnd(32 * W + 1), n = (int)rnd(32 * W + 1);
        if (t % 3 == 0) n = m;  // related pairs: mutate a copy
        for (int i = 0; i < m; ++i) a[i] = (uint8_t)rnd(K);
        for (int i = 0; i < n; ++i) b[i] = (t % 3 == 0 && i < m && rnd(8)) ? a[i] : (uint8_t)rnd(K);
        int e = wf(a, m, b, n), g = myers<W>(a, m, b, n, K);
        if (e != g) {
            if (bad < 5) std::printf("W=%d m=%d n=%d wf=%d myers=%d\n", W, m, n, e, g);
            ++bad;
        }
    }
    return bad;
}

int main()
{
    int bad = 0;
    const int Ks[] = {1, 2, 4, 20, 32};
    for (int K : Ks) {
        bad += check_w<1>(20000, K);
        bad += check_w<2>(5000, K);
        bad += check_w<3>(3000, K);
        bad += check_w<5>(2000, K);
        bad += check_w<8>(500, K);
    }
    // packed Hamming
    for (int t = 0; t < 100000; ++t) {
        uint8_t x[4], y[4];
        for (int i = 0; i < 4; ++i) { x[i] = (uint8_t)rnd(256); y[i] = rnd(2) ? x[i] : (uint8_t)rnd(256); }
        uint32_t a, b;
        std::memcpy(&a, x, 4); std::memcpy(&b, y, 4);
        uint32_t e = 0;
        for (int i = 0; i < 4; ++i) e += x[i] != y[i];
        if (idb_mismatch8(a, b) != e) ++bad;
        uint32_t a2 = 0, b2 = 0, e2 = 0;
        for (int i = 0; i < 16; ++i) {
            uint32_t c = rnd(4), d = rnd(2) ? c : rnd(4);
            a2 |= c << (2 * i); b2 |= d << (2 * i); e2 += c != d;
        }
        if (idb_mismatch2(a2, b2) != e2) ++bad;
        uint32_t a5 = 0, b5 = 0, e5 = 0;
        for (int i = 0; i < 6; ++i) {
            uint32_t c = rnd(32), d = rnd(2) ? c : rnd(32);
            a5 |= c << (5 * i); b5 |= d << (5 * i); e5 += c != d;
        }
        if (idb_mismatch5(a5, b5) != e5) ++bad;
    }
    std::printf("%s mismatches=%d\n", bad ? "FAIL" : "OK", bad);
    return bad ? 1 : 0;
}
