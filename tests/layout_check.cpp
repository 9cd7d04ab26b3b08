// Host check of the lattice-state layout helpers (IonLayout) and of the
// compensated FP32 value format (pk_pack / pk_unpack) in
// php_fdtd_amd/csrc/phpfdtd_internal.h.  Built and run by test_layout.py;
// prints "ok" or the first violated property and exits non-zero.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include <string.h>

#include <random>
#include <set>
#include <vector>

#include "../php_fdtd_amd/csrc/phpfdtd_internal.h"

using namespace phpfdtd;

#define CHECK(c, ...)                          \
    do {                                       \
        if (!(c)) {                            \
            printf("FAIL %s: ", #c);           \
            printf(__VA_ARGS__);               \
            printf("\n");                      \
            exit(1);                           \
        }                                      \
    } while (0)

static void check_layout(int natom, bool packed) {
    const IonLayout l = ion_layout(natom, packed);
    const int n3 = 3 * natom, w = packed ? 4 : 2;
    const int64_t es = packed ? 4 : 8;  // bytes per element
    const int64_t tile = l.tile_r(), tilev = l.tile_v();
    CHECK(l.n3p() % w == 0 && l.n3p() >= n3 && l.n3p() < n3 + w, "natom %d", natom);
    CHECK(l.ng() * w == l.n3p(), "natom %d", natom);
    // v's last codes share r's last compensation row exactly when that row holds 1-4 groups
    const int tail = l.ng() % 8;
    const bool shared = packed && tail >= 1 && tail <= 4;
    CHECK(l.shared() == shared, "natom %d", natom);
    if (packed) {
        // words of ng groups, then ceil(ng/8) rows of 4-bit compensation codes, 1 KiB each;
        // v's tile has one row fewer when shared
        const int nb = (l.ng() + 7) / 8;
        CHECK(tile * es == 1024LL * (l.ng() + nb), "natom %d tile %lld", natom, (long long)tile);
        CHECK(tilev * es == 1024LL * (l.ng() + nb - (shared ? 1 : 0)), "natom %d v tile %lld", natom, (long long)tilev);
    } else {
        CHECK(tile * es == 64LL * 8 * l.n3p() && tilev == tile, "natom %d tile %lld", natom, (long long)tile);
    }
    std::set<int64_t> words, codes, vwords, vcodes;
    for (int t = 0; t < 2; ++t)
        for (int lane = 0; lane < 64; ++lane)
            for (int q = 0; q < l.n3p(); ++q) {
                const int64_t s = 64LL * t + lane;
                const int64_t ri = l.r_index(s, q);
                CHECK(ri >= t * tile && ri < (t + 1) * tile, "natom %d slot %lld dof %d", natom, (long long)s, q);
                const int64_t vi = l.v_index(s, q);
                CHECK(vi - t * tilev == ri - t * tile, "natom %d", natom);  // same place in the tile
                CHECK(words.insert(ri).second, "natom %d: element %lld used twice", natom, (long long)ri);
                CHECK(vwords.insert(vi).second, "natom %d: v element %lld used twice", natom, (long long)vi);
                // a lane's w dofs of one group are one 16-byte vector, the 64 lanes' vectors one 1 KiB row
                CHECK((ri - t * tile) * es == 1024LL * (q / w) + 16LL * lane + es * (q % w), "natom %d slot %lld dof %d",
                      natom, (long long)s, q);
                int64_t s2 = -1;
                int q2 = -1;
                const bool live = l.r_decode(ri, 1 << 20, &s2, &q2);
                CHECK(live == (q < n3), "natom %d dof %d: decode says %d", natom, q, (int)live);
                if (live) CHECK(s2 == s && q2 == q, "natom %d: decode(%lld) = (%lld, %d)", natom, (long long)ri,
                                (long long)s2, q2);
                CHECK(!l.r_decode(ri, s, &s2, &q2), "natom %d: slot %lld past nslot decoded", natom, (long long)s);
                if (packed) {
                    const int64_t ci = l.c_index(s, q);  // nibble index
                    const int64_t cw = ci / 8;            // its dword
                    // compensation codes sit in the rows after the ng word rows of the same tile
                    CHECK(cw >= t * tile + (int64_t)kPkWords * l.ng() && cw < (t + 1) * tile,
                          "natom %d slot %lld dof %d code %lld", natom, (long long)s, q, (long long)ci);
                    // one 16-byte vector per lane and row: the lane's codes of groups 8b .. 8b+7, two
                    // groups per dword, the group's four codes in one 16-bit half
                    const int j = q / 4;
                    CHECK((cw - t * tile - (int64_t)kPkWords * l.ng()) * 4 == 1024LL * (j / 8) + 16LL * lane + 4 * ((j % 8) / 2) &&
                              ci % 8 == 4 * (j % 2) + q % 4,
                          "natom %d slot %lld dof %d", natom, (long long)s, q);
                    CHECK(codes.insert(ci).second, "natom %d: code %lld used twice", natom, (long long)ci);
                    // v's code: the same place in v's tile, or, for a group of r's shared last row,
                    // the same bits of dword 2 + (j % 8) / 2 of that row in r's tile
                    bool in_r = false;
                    const int64_t cv = l.cv_index(s, q, &in_r);
                    CHECK(in_r == (shared && j / 8 == (l.ng() + 7) / 8 - 1), "natom %d dof %d", natom, q);
                    if (in_r) {
                        CHECK(cv == ci + 16, "natom %d slot %lld dof %d", natom, (long long)s, q);
                        CHECK(codes.insert(cv).second, "natom %d: v code %lld on an r code", natom, (long long)cv);
                    } else {
                        CHECK(cv - t * tilev * 8 == ci - t * tile * 8, "natom %d slot %lld dof %d", natom, (long long)s, q);
                        CHECK(cv / 8 >= t * tilev + (int64_t)kPkWords * l.ng() && cv / 8 < (t + 1) * tilev,
                              "natom %d slot %lld dof %d v code %lld", natom, (long long)s, q, (long long)cv);
                        CHECK(vcodes.insert(cv).second, "natom %d: v code %lld used twice", natom, (long long)cv);
                    }
                }
            }
    if (packed) {
        // every element of the compensation rows decodes as no dof
        for (int64_t e = (int64_t)kPkWords * l.ng(); e < tile; ++e) {
            int64_t s2;
            int q2;
            CHECK(!l.r_decode(e, 1 << 20, &s2, &q2), "natom %d: compensation element %lld decoded", natom, (long long)e);
        }
        CHECK((int64_t)words.size() * 8 + (int64_t)codes.size() <= 2 * tile * 8, "natom %d", natom);
    } else {
        CHECK((int64_t)words.size() == 2 * tile, "natom %d: FP64 tiles not covered", natom);
    }
}

// band_tile: over a launch of a multiple of 8 g work-groups every tile has
// exactly one work-group, and the g tiles of a group run on one XCD, the
// groups going to the XCDs in turn
static void check_band_tile() {
    for (int g = 1; g <= 9; ++g)
        for (int n = 1; n <= 700; n += (n < 40 ? 1 : 37)) {
            const int pad = (n + 8 * g - 1) / (8 * g) * (8 * g);
            std::vector<int> seen(n, -1);
            for (int b = 0; b < pad; ++b) {
                const int t = band_tile(b, n, g);
                CHECK(t >= -1 && t < n, "g %d n %d b %d -> %d", g, n, b, t);
                if (t < 0) continue;
                CHECK(seen[t] < 0, "g %d n %d: tile %d twice", g, n, t);
                seen[t] = b;
                CHECK(b % 8 == (t / g) % 8, "g %d n %d: tile %d on XCD %d", g, n, t, b % 8);
            }
            for (int t = 0; t < n; ++t) CHECK(seen[t] >= 0, "g %d n %d: tile %d never run", g, n, t);
        }
}

static void check_values() {
    std::mt19937_64 g(5);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    std::uniform_int_distribution<int> ex(-140, 120);
    // zero is exact
    float hi;
    int lo;
    pk_pack(0.0, &hi, &lo);
    CHECK(hi == 0.f && lo == 0 && pk_unpack(hi, lo) == 0.0, "zero");
    for (int n = 0; n < 200000; ++n) {
        const double x = ldexp(u(g), ex(g));
        pk_pack(x, &hi, &lo);
        CHECK(lo >= -7 && lo <= 7, "x %.17g lo %d", x, lo);
        CHECK(hi == (float)x, "x %.17g: word is not fl32(x)", x);
        const double y = pk_unpack(hi, lo);
        // within half a unit (2^-4 ulp of the word) where the correction fits
        // the code, one unit where a +-8 is clamped to +-7
        const int e = pk_exp(hi);
        const double unit = ldexp(1.0, e - 154);
        const bool clamped = lo == 7 || lo == -7;
        CHECK(fabs(y - x) <= (clamped ? 1.0 : 0.5) * unit, "x %.17g y %.17g unit %g lo %d", x, y, unit, lo);
        // relative, normal words: 28 significant bits (2^-28), 2^-27 when clamped
        if (fabs(x) >= ldexp(1.0, -125))
            CHECK(fabs(y - x) <= ldexp(fabs(x), clamped ? -27 : -28), "x %.17g y %.17g", x, y);
        // idempotent: a stored value re-packs to the same code
        float hi2;
        int lo2;
        pk_pack(y, &hi2, &lo2);
        CHECK(hi2 == hi && lo2 == lo, "x %.17g: (%a, %d) -> (%a, %d)", x, hi, lo, hi2, lo2);
    }
    // just under the midpoint between 1 and its FP32 successor: the word is 1,
    // the correction rounds to 8 and is clamped to 7, an error of 2^-27
    const double x = 1.0 + ldexp(1.0, -24) - ldexp(1.0, -60);
    pk_pack(x, &hi, &lo);
    CHECK(hi == 1.f && lo == 7, "clamp x %.17g: (%a, %d)", x, hi, lo);
    CHECK(fabs(pk_unpack(hi, lo) - x) <= ldexp(fabs(x), -27), "clamp x %.17g", x);
    // the codes at nibble indices: sign extension and replacement leave the neighbour alone
    uint8_t buf[4] = {0, 0, 0, 0};
    for (int c = -8; c <= 7; ++c)
        for (int n = 0; n < 8; ++n) {
            pk_code_set(buf, n, c);
            CHECK(pk_code_at(buf, n) == c, "code %d at %d", c, n);
            uint32_t dw;
            memcpy(&dw, buf, 4);
            CHECK(pk_code_of(dw, 4 * n) == c, "dword code %d at %d", c, n);
            CHECK(pk_code_of(pk_with_code(dw, 4 * n, -c - 1), 4 * n) == -c - 1, "with_code %d at %d", c, n);
            CHECK(pk_code_at(buf, n ^ 1) == pk_code_of(dw, 4 * (n ^ 1)), "neighbour of %d", n);
        }
}

int main() {
    for (int natom = 1; natom <= kMaxAtom; ++natom) {
        check_layout(natom, false);
        check_layout(natom, true);
    }
    check_values();
    check_band_tile();
    printf("ok\n");
    return 0;
}
