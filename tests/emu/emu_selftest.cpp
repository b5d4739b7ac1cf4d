// Kernels with a planted fault each, to show that the CPU emulation reports
// that kind of fault (test infrastructure).
//
//   emu_selftest CASE ORDER
//
// CASE: clean (a correct kernel: exit 0), barrier (a barrier that only the
// first wave reaches, while the others wait at the next one), halfwave (a lane read by half of a wave), race (a
// shared-memory value read without a barrier after thread 0 writes it: the
// result depends on the lane order), oob (a global write one element past
// the end of its buffer: the address sanitizer aborts).  ORDER as in
// emu::launch.  Prints the per-thread result checksum on success; exit 2 on
// an emulation error.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "emu_runtime.hpp"

namespace {

int g_case = 0;

void k_selftest(int *out) {
    __shared__ int flag;
    __shared__ int vals[256];
    const int tid = (int)threadIdx.x, lane = tid & 63;
    vals[tid] = tid * 3;
    if (tid == 0) flag = 0;
    __syncthreads();
    switch (g_case) {
    case 0: {  // clean: lane reads, a DPP row rotation, a ballot, barriers
        const int v = __builtin_amdgcn_readlane(vals[tid], 5);
        const int r = __builtin_amdgcn_update_dpp(tid, tid, 0x121, 0xF, 0xF, false);
        const unsigned long long b = __ballot(lane & 1);
        __syncthreads();
        out[tid] = v + r + (int)(b & 0xFF) + vals[255 - tid];
        break;
    }
    case 1:  // a barrier only the first wave reaches, before one that every wave reaches
        if (tid < 64) __syncthreads();
        __syncthreads();
        out[tid] = vals[tid];
        break;
    case 2: {  // collective reached by half of a wave
        int v = 0;
        if (lane < 32) v = __builtin_amdgcn_readlane(tid, 0);
        out[tid] = v;
        break;
    }
    case 3:  // shared value read without the barrier after it is written
        if (tid == 0) flag = 7;
        out[tid] = flag;
        break;
    case 4:  // one element past the end of `out`
        out[tid + 1] = tid;
        break;
    }
}

void body(void *p) { k_selftest(static_cast<int *>(p)); }

}  // namespace

int main(int argc, char **argv) {
    if (argc != 3) return 1;
    const char *names[] = {"clean", "barrier", "halfwave", "race", "oob"};
    g_case = -1;
    for (int i = 0; i < 5; ++i)
        if (!std::strcmp(argv[1], names[i])) g_case = i;
    if (g_case < 0) return 1;
    int *out = static_cast<int *>(std::malloc(256 * sizeof(int)));
    std::string err;
    emu::Stats st;
    if (emu::launch(0, 1, 256, std::atoi(argv[2]), 7u, body, out, &err, &st) != 0) {
        std::fprintf(stderr, "emulation error: %s\n", err.c_str());
        return 2;
    }
    unsigned long long sum = 0;
    for (int i = 0; i < 256; ++i) sum = sum * 31u + (unsigned)out[i];
    std::printf("%llu\n", sum);
    std::free(out);
    return 0;
}
