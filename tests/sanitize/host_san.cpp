// host_san.cpp -- the library's host-only code under AddressSanitizer and
// UndefinedBehaviorSanitizer: tree and list construction (tree.cpp, the
// part of the path that must match the reference bit for bit) and the
// reference's cereal vector files (vecio.cpp).  Built by the Makefile beside
// this file with the host compiler, run by tests/test_sanitizers.py.
//
// Checks on every tree: the permutation is one; every level's nodes
// partition the points into contiguous ranges nested in their parent's;
// every point lies in its node's box; leaves keep ascending ids; lists
// hold no self entry, only nodes of the level, and interaction lists are
// symmetric.  Vector files: round
// trips, and truncated, lying or foreign files refused without reading past
// their end.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../hodlrnd_amd/csrc/internal.h"

static int fails = 0;
#define CHECK(c, ...)                                              \
    do {                                                           \
        if (!(c)) {                                                \
            std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
            std::fprintf(stderr, __VA_ARGS__);                     \
            std::fputc('\n', stderr);                              \
            fails++;                                               \
        }                                                          \
    } while (0)

static uint64_t sm_state;
static double uniform() {  // SplitMix64 -> [-1, 1)
    uint64_t z = (sm_state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    return -1.0 + 2.0 * (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

static void check_tree(const hodlr_tree &t, const char *name) {
    const int d = t.ndim;
    std::vector<char> seen(t.n, 0);
    for (int64_t q = 0; q < t.n; q++) {
        const int64_t id = t.perm[q];
        CHECK(id >= 0 && id < t.n && !seen[id], "%s: perm is not a permutation", name);
        if (id >= 0 && id < t.n) seen[id] = 1;
    }
    for (int l = 0; l < t.nlevels; l++) {
        int64_t pos = 0;
        for (int64_t j = 0; j < t.level_size(l); j++) {
            const int64_t g = t.node(l, j);
            CHECK(t.start[g] == pos, "%s: level %d node %lld does not start where the last ended", name, l,
                  (long long)j);
            pos += t.count[g];
            if (l > 0) {
                const int64_t par = t.node(l - 1, j >> d);
                CHECK(t.start[g] >= t.start[par] && t.start[g] + t.count[g] <= t.start[par] + t.count[par],
                      "%s: child range outside the parent's", name);
            }
            for (int64_t a = t.start[g]; a < t.start[g] + t.count[g]; a++) {
                const double *x = &t.pts[t.perm[a] * d];
                for (int k = 0; k < d; k++)
                    CHECK(x[k] >= t.lo[g * d + k] && x[k] <= t.hi[g * d + k], "%s: point outside its box", name);
                // leaves keep ascending point ids (the stable splits); an
                // inner node's range is its children's ranges in turn
                if (l == t.nlevels - 1 && a > t.start[g])
                    CHECK(t.perm[a] > t.perm[a - 1], "%s: leaf ids not ascending", name);
            }
            for (int which = 0; which < 2; which++) {
                const auto &ptr = which ? t.il_ptr : t.nbr_ptr;
                const auto &v = which ? t.il : t.nbr;
                for (int64_t a = ptr[g]; a < ptr[g + 1]; a++) {
                    const int64_t o = v[a];
                    CHECK(o >= 0 && o < t.level_size(l) && o != j, "%s: bad list entry", name);
                    if (which == 1 && o >= 0 && o < t.level_size(l)) {
                        const int64_t go = t.node(l, o);
                        bool back = false;
                        for (int64_t b = t.il_ptr[go]; b < t.il_ptr[go + 1]; b++) back = back || t.il[b] == j;
                        CHECK(back, "%s: interaction list not symmetric", name);
                    }
                }
            }
        }
        CHECK(pos == t.n, "%s: level %d covers %lld of %lld points", name, l, (long long)pos, (long long)t.n);
    }
}

static void tree_case(const char *name, int d, int64_t n, int kind, int nmax, int dprime) {
    hodlr_tree t;
    t.ndim = d;
    t.nmax = nmax;
    t.dprime = dprime;
    t.lo.assign(d, -1.0);
    t.hi.assign(d, 1.0);
    if (kind == 1) {  // tensor grid: coordinates on the split planes
        int m = (int)std::floor(std::pow((double)n, 1.0 / d) + 1e-9);
        n = 1;
        for (int k = 0; k < d; k++) n *= m;
        t.pts.resize(n * d);
        for (int64_t i = 0; i < n; i++) {
            int64_t r = i;
            for (int k = d - 1; k >= 0; k--) {
                t.pts[i * d + k] = -1.0 + 2.0 * (double)(r % m) / (m - 1);
                r /= m;
            }
        }
    } else {
        t.pts.resize(n * d);
        for (auto &x : t.pts) x = kind == 2 ? -1.0 + 0.25 * (uniform() + 1.0) : uniform();
    }
    t.n = n;
    hodlr::build_tree(t);
    check_tree(t, name);
    std::fprintf(stderr, "ok %-24s n=%-7lld levels=%d nodes=%lld\n", name, (long long)n, t.nlevels,
                 (long long)t.level_off.back());
}

static void coincident_points_refused() {
    hodlr_tree t;
    t.ndim = 2;
    t.nmax = 4;
    t.lo.assign(2, -1.0);
    t.hi.assign(2, 1.0);
    t.n = 9;
    t.pts.assign(18, 0.25);
    bool threw = false;
    try {
        hodlr::build_tree(t);
    } catch (const hodlr::ArgError &e) {
        threw = e.code == HODLR_E_ARG;
    }
    CHECK(threw, "coincident points were not refused");
    std::fprintf(stderr, "ok coincident points refused\n");
}

static void write_bytes(const char *path, const void *p, size_t nb) {
    FILE *f = std::fopen(path, "wb");
    if (!f) return;
    if (nb) std::fwrite(p, 1, nb, f);
    std::fclose(f);
}

static void vector_files(const char *dir) {
    char path[4096];
    std::snprintf(path, sizeof(path), "%s/v.bin", dir);
    std::vector<double> v(1000), w(1000);
    for (auto &x : v) x = uniform();
    CHECK(hodlr_vec_save(path, 1000, v.data()) == HODLR_OK, "save");
    int64_t n = -1;
    CHECK(hodlr_vec_load(path, 0, nullptr, &n) == HODLR_OK && n == 1000, "count only");
    CHECK(hodlr_vec_load(path, 1000, w.data(), &n) == HODLR_OK && std::memcmp(v.data(), w.data(), 8000) == 0,
          "round trip");
    CHECK(hodlr_vec_load(path, 999, w.data(), &n) == HODLR_E_ARG, "capacity too small accepted");
    CHECK(hodlr_vec_save(path, 0, nullptr) == HODLR_OK, "empty save");
    CHECK(hodlr_vec_load(path, 0, w.data(), &n) == HODLR_OK && n == 0, "empty load");
    // truncated: header promises 1000 values, the file holds 10
    std::vector<unsigned char> b(9 + 80);
    b[0] = 1;
    const uint64_t cnt = 1000;
    std::memcpy(&b[1], &cnt, 8);
    write_bytes(path, b.data(), b.size());
    CHECK(hodlr_vec_load(path, 1000, w.data(), &n) == HODLR_E_ARG, "truncated file accepted");
    CHECK(hodlr_vec_load(path, 0, nullptr, &n) == HODLR_E_ARG, "truncated file's count reported");
    // a count near 2^64
    const uint64_t huge = ~(uint64_t)0;
    std::memcpy(&b[1], &huge, 8);
    write_bytes(path, b.data(), b.size());
    CHECK(hodlr_vec_load(path, 1000, w.data(), &n) == HODLR_E_ARG, "huge count accepted");
    // not a little-endian portable archive
    b[0] = 0;
    write_bytes(path, b.data(), b.size());
    CHECK(hodlr_vec_load(path, 1000, w.data(), &n) == HODLR_E_ARG, "foreign archive accepted");
    write_bytes(path, b.data(), 4);  // shorter than the header
    CHECK(hodlr_vec_load(path, 1000, w.data(), &n) == HODLR_E_ARG, "short header accepted");
    std::snprintf(path, sizeof(path), "%s/missing/v.bin", dir);
    CHECK(hodlr_vec_load(path, 1000, w.data(), &n) == HODLR_E_ARG, "missing file");
    std::fprintf(stderr, "ok vector files\n");
}

int main(int argc, char **argv) {
    sm_state = 12345;
    tree_case("1d random", 1, 5000, 0, 16, 0);
    tree_case("2d random", 2, 20000, 0, 32, 0);
    tree_case("2d grid d'=-1", 2, 10000, 1, 32, -1);
    tree_case("3d random", 3, 40000, 0, 64, 0);
    tree_case("3d grid", 3, 27000, 1, 64, 0);
    tree_case("3d clustered", 3, 4000, 2, 32, 1);
    tree_case("4d random d'=3", 4, 30000, 0, 64, 3);
    tree_case("4d grid", 4, 4096, 1, 32, 0);
    tree_case("5d random", 5, 3000, 0, 64, 0);
    tree_case("depth zero", 3, 40, 0, 64, 0);
    tree_case("empty", 2, 0, 0, 8, 0);
    coincident_points_refused();
    vector_files(argc > 1 ? argv[1] : "/tmp");
    if (fails) {
        std::fprintf(stderr, "%d checks failed\n", fails);
        return 1;
    }
    std::fprintf(stderr, "host sanitizer run clean\n");
    return 0;
}
