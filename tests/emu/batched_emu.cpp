// Runs the batched solver's kernels (rulinalg_amd/csrc/batched_device.hpp,
// the same source the library compiles for gfx950) on the CPU under the
// emulation of tests/emu (test infrastructure).
//
//   batched_emu IN OUT
//
// IN (little-endian): int32 mode (0 plain k_batch_f64, 1 mixed
// k_batch_mixed), count, block (the workgroup index the run takes, which
// selects the mixed workspace slot), grid, order (0 ascending, 1 descending,
// 2 shuffled), seed, max_iter; float64 tol; then count x int32 {kind, n,
// nrhs, nb}; then for each item its A (n x n) and its right-hand sides
// (nrhs x n), row-major float64.
// OUT: int32 status[count], int32 fallbacks; int64 barriers, collectives;
// then for each item A (n x n), B (nrhs x n), ipiv (n, int64).
//
// Every buffer is its own heap allocation of exactly its size, so that the
// address sanitizer reports any access outside it.  Exit status 0 on
// success, 2 on an emulation error (printed), 1 on bad input.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "emu_runtime.hpp"
#include "../../rulinalg_amd/csrc/batched_device.hpp"

namespace {

struct Args {
    int mode, count, grid, max_iter;
    double tol;
    rlg_batch_item *items;
    int *order;
    unsigned char *nbs;
    unsigned *counter;
    char *ws;
    unsigned *fallbacks;
};

void body(void *p) {
    Args *a = static_cast<Args *>(p);
    if (a->mode == 0)
        rlg::k_batch_f64(a->items, a->order, a->nbs, a->count, a->counter);
    else
        rlg::k_batch_mixed(a->items, a->order, a->nbs, a->count, a->counter, a->ws, a->max_iter, a->tol,
                           a->fallbacks);
}

bool rd(FILE *f, void *p, size_t n) { return std::fread(p, 1, n, f) == n; }

}  // namespace

int main(int argc, char **argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: batched_emu IN OUT\n");
        return 1;
    }
    FILE *f = std::fopen(argv[1], "rb");
    if (!f) return 1;
    int hdr[7];
    double tol;
    if (!rd(f, hdr, sizeof hdr) || !rd(f, &tol, 8)) return 1;
    const int mode = hdr[0], count = hdr[1], block = hdr[2], grid = hdr[3], order = hdr[4], seed = hdr[5],
              max_iter = hdr[6];
    if (count < 1 || grid < 1 || block < 0 || block >= grid) return 1;
    std::vector<int> meta(4 * (size_t)count);
    if (!rd(f, meta.data(), meta.size() * 4)) return 1;
    // device-side arrays, each allocated at its exact size
    auto *items = static_cast<rlg_batch_item *>(std::malloc(sizeof(rlg_batch_item) * count));
    auto *ord = static_cast<int *>(std::malloc(sizeof(int) * count));
    auto *nbs = static_cast<unsigned char *>(std::malloc(count));
    auto *counter = static_cast<unsigned *>(std::calloc(2, sizeof(unsigned)));
    auto *fallbacks = static_cast<unsigned *>(std::calloc(1, sizeof(unsigned)));
    for (int i = 0; i < count; ++i) {
        const int kind = meta[4 * i], n = meta[4 * i + 1], nrhs = meta[4 * i + 2], nb = meta[4 * i + 3];
        if (n < 1 || n > RLG_BATCH_MAX_N || nrhs < 1 || nrhs > RLG_BATCH_MAX_NRHS) return 1;
        auto *a = static_cast<double *>(std::malloc(8 * (size_t)n * n));
        auto *b = static_cast<double *>(std::malloc(8 * (size_t)nrhs * n));
        auto *ip = static_cast<int64_t *>(std::malloc(8 * (size_t)n));
        if (!rd(f, a, 8 * (size_t)n * n) || !rd(f, b, 8 * (size_t)nrhs * n)) return 1;
        for (int j = 0; j < n; ++j) ip[j] = -1;
        items[i] = rlg_batch_item{kind, n, nrhs, -1, a, b, ip};
        ord[i] = i;
        nbs[i] = (unsigned char)nb;
    }
    std::fclose(f);
    char *ws = nullptr;
    if (mode == 1) ws = static_cast<char *>(std::malloc(rlg::SLOT_BYTES * (size_t)grid));
    Args args{mode, count, grid, max_iter, tol, items, ord, nbs, counter, ws, fallbacks};
    std::string err;
    emu::Stats st;
    if (emu::launch(block, grid, rlg::BTH, order, (unsigned)seed, body, &args, &err, &st) != 0) {
        std::fprintf(stderr, "emulation error: %s\n", err.c_str());
        return 2;
    }
    FILE *o = std::fopen(argv[2], "wb");
    if (!o) return 1;
    for (int i = 0; i < count; ++i) std::fwrite(&items[i].status, 4, 1, o);
    const int fb = (int)*fallbacks;
    std::fwrite(&fb, 4, 1, o);
    const long long sb = (long long)st.barriers, sc = (long long)st.collectives;
    std::fwrite(&sb, 8, 1, o);
    std::fwrite(&sc, 8, 1, o);
    for (int i = 0; i < count; ++i) {
        const rlg_batch_item &it = items[i];
        std::fwrite(it.a, 8, (size_t)it.n * it.n, o);
        std::fwrite(it.b, 8, (size_t)it.nrhs * it.n, o);
        std::fwrite(it.ipiv, 8, (size_t)it.n, o);
        std::free(it.a);
        std::free(it.b);
        std::free(it.ipiv);
    }
    std::fclose(o);
    std::free(ws);
    std::free(items);
    std::free(ord);
    std::free(nbs);
    std::free(counter);
    std::free(fallbacks);
    return 0;
}
