// CPU check of the filter layer's plane copies (csrc/host_copy.h): random row widths, row counts,
// strides and source/destination misalignments, against row-by-row memcpy; bytes outside the
// copied rows must stay untouched.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "host_copy.h"

int main() {
    std::mt19937 rng(1);
    for (int it = 0; it < 20000; ++it) {
        const size_t rb = rng() % 9000, rows = 1 + rng() % 5;
        const size_t so = rng() % 64, doff = rng() % 64;
        const ptrdiff_t ss = ptrdiff_t(rb + rng() % 100), ds = ptrdiff_t(rb + rng() % 100);
        std::vector<char> s(so + size_t(ss) * rows + 64), d(doff + size_t(ds) * rows + 64, 7);
        for (auto &c : s) c = char(rng());
        std::vector<char> ref = d;
        for (size_t y = 0; y < rows; ++y) std::memcpy(ref.data() + doff + y * ds, s.data() + so + y * ss, rb);
        vsz_host::copy_plane(d.data() + doff, ds, s.data() + so, ss, rb, int(rows));
        if (d != ref) {
            std::printf("mismatch: iteration %d, row %zu bytes x %zu rows, src +%zu, dst +%zu\n", it, rb, rows, so, doff);
            return 1;
        }
    }
    std::printf("host copy: ok\n");
    return 0;
}
