// Host check of the 3D sweep mapping (deepgpr_amd/csrc/dgx_sweep.h): every
// cell of the field grid is advanced by exactly one thread of exactly one
// block, cell_map_block inverts the mapping, the box cells meet the
// conditions that let the kernels skip every face test, and the
// multiply-and-shift divisions equal integer division for q < 2^31.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "dgx_sweep.h"

using namespace dgx;

static int failures = 0;
#define CHECK(cond, ...)                       \
    do {                                       \
        if (!(cond)) {                         \
            if (failures < 20) {               \
                std::printf("FAIL: " __VA_ARGS__); \
                std::printf("\n");             \
            }                                  \
            ++failures;                        \
        }                                      \
    } while (0)

// Conditions of a plain box cell on one axis, restated from the kernels'
// helpers: outside the E and H CPML face regions, unclamped differences of
// radius R, every component updated, full-width gathers.
static bool axis_plain(int c, int n, int plo, int phi, int R)
{
    if (plo > 0 && c >= 0 && c <= plo) return false;                 // E: 1..p, H: 0..p-1
    if (phi > 0 && c >= n - 1 - phi && c <= n - 2) return false;     // both: n-1-p .. n-2
    if (c < R || c > n - 1 - R) return false;
    if (c < 1 || c > n - 2) return false;
    if (c < 2 * R - 1 || c > n - 2 * R) return false;
    return true;
}

static void run_case(int n0, int n1, int n2, const int pml[6], int R, bool use_box, bool ilv)
{
    const int N[3] = {n0, n1, n2};
    CellMap cm = make_cell_map(N, pml, R, use_box);
    if (ilv) interleave(cm);
    std::vector<unsigned char> seen((size_t)n0 * n1 * n2, 0);
    long long box_cells = 0;
    for (int b = 0; b < cm.per_shot; ++b)
        for (int t = 0; t < 256; ++t) {
            int i = -1, j = -1, k = -1, idx = -1;
            const int part = sweep_part(cm, b, idx);
            if (part < 0) continue;
            const bool boxed = part == 1;
            if (boxed) CHECK(idx % 8 == b % 8, "box block %d of launch block %d is on another XCD", idx, b);
            const bool ok = boxed ? box_cell(cm, idx, t, i, j, k) : shell_cell(cm, idx, t, i, j, k);
            if (!ok) continue;
            CHECK(i >= 0 && i < n0 && j >= 0 && j < n1 && k >= 0 && k < n2, "cell (%d,%d,%d) of block %d outside %dx%dx%d",
                  i, j, k, b, n0, n1, n2);
            if (!(i >= 0 && i < n0 && j >= 0 && j < n1 && k >= 0 && k < n2)) continue;
            unsigned char& s = seen[((size_t)i * n1 + j) * n2 + k];
            CHECK(s == 0, "cell (%d,%d,%d) advanced twice (block %d)", i, j, k, b);
            s = 1;
            CHECK(cell_map_block(cm, i, j, k) == b, "cell_map_block(%d,%d,%d) = %lld, thread of block %d", i, j, k,
                  cell_map_block(cm, i, j, k), b);
            if (boxed) {
                ++box_cells;
                CHECK(axis_plain(i, n0, pml[0], pml[1], R) && axis_plain(j, n1, pml[2], pml[3], R) &&
                          axis_plain(k, n2, pml[4], pml[5], R),
                      "box cell (%d,%d,%d) is not plain (grid %dx%dx%d, R %d)", i, j, k, n0, n1, n2, R);
            }
        }
    long long missing = 0;
    for (unsigned char s : seen) missing += s == 0;
    CHECK(missing == 0, "%lld cells of %dx%dx%d not advanced (R %d, box %d)", missing, n0, n1, n2, R, (int)use_box);
    CHECK(cm.shell_end % 8 == 0 && cm.per_shot % 8 == 0, "block counts not multiples of 8");
    std::printf("grid %dx%dx%d R %d box %d interleave %d: %d blocks per shot, %lld box cells\n", n0, n1, n2, R,
                (int)use_box, (int)ilv, cm.per_shot, box_cells);
}

static void check_div()
{
    unsigned seed = 12345;
    auto rnd = [&]() { seed = seed * 1664525u + 1013904223u; return seed; };
    const unsigned edge_q[] = {0u, 1u, 2u, 255u, 256u, 65535u, 65536u, (1u << 24) - 1, 1u << 24,
                               (1u << 31) - 2, (1u << 31) - 1};
    for (int trial = 0; trial < 4000; ++trial) {
        unsigned d = trial < 64 ? (unsigned)trial + 1 : 1 + rnd() % (trial % 2 ? 5000u : (1u << 30));
        if (trial % 7 == 0) d = 1u << (rnd() % 31);
        const Div v = make_div(d);
        for (unsigned q : edge_q) CHECK(div_q(q, v) == q / d, "div_q(%u, %u) = %u", q, d, div_q(q, v));
        for (int r = 0; r < 300; ++r) {
            const unsigned q = rnd() & 0x7fffffffu;
            CHECK(div_q(q, v) == q / d, "div_q(%u, %u) = %u", q, d, div_q(q, v));
            const unsigned near = (q / d) * d;   // multiples of d and their neighbours
            for (unsigned x : {near, near - 1, near + d - 1})
                if (x < (1u << 31)) CHECK(div_q(x, v) == x / d, "div_q(%u, %u) = %u", x, d, div_q(x, v));
        }
    }
}

int main()
{
    check_div();
    const int c3[6] = {16, 16, 16, 16, 16, 16};
    const int none[6] = {0, 0, 0, 0, 0, 0};
    const int uneven[6] = {5, 4, 3, 6, 4, 5};
    const int thin[6] = {1, 2, 2, 1, 1, 2};
    const int t1[6] = {5, 4, 3, 6, 4, 5};   // grids of the GPU sweep-map test
    const int t2[6] = {4, 5, 3, 4, 6, 5};
    for (int ilv = 0; ilv < 2; ++ilv) {
        run_case(289, 289, 225, c3, 2, true, ilv);    // the 3D benchmark model, order 4
        run_case(129, 129, 129, c3, 2, true, ilv);
        run_case(129, 129, 129, c3, 4, true, ilv);
        run_case(97, 81, 113, none, 4, true, ilv);   // no CPML, order 8
        run_case(97, 81, 113, none, 1, true, ilv);
        run_case(61, 72, 90, uneven, 2, true, ilv);
        run_case(61, 72, 90, uneven, 1, false, ilv);
        run_case(30, 26, 34, uneven, 2, true, ilv);   // too small for a box
        run_case(7, 8, 6, thin, 1, true, ilv);
        run_case(40, 33, 2, none, 2, true, ilv);      // one z-plane of cells
        for (int R = 1; R <= 4; R *= 2) run_case(60, 56, 70, t1, R, true, ilv);
        run_case(56, 52, 70, none, 2, true, ilv);
        run_case(60, 40, 270, t2, 2, true, ilv);
    }
    if (failures) {
        std::printf("%d checks failed\n", failures);
        return 1;
    }
    std::printf("all checks passed\n");
    return 0;
}
