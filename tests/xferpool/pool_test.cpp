// Host test of pk_xfer::CopyPool, the threaded memcpy behind the library's
// host <-> device transfers: many copies of odd sizes and offsets, back to
// back, each checked byte for byte; then stop and start again.  Prints "ok".
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../piernik_amd/csrc/pk_xfer.hpp"

int main()
{
    pk_xfer::CopyPool pool;
    pool.start(5);
    const size_t cap = (size_t)40 << 20;
    std::vector<unsigned char> src(cap), dst(cap);
    unsigned seed = 12345u;
    for (size_t i = 0; i < cap; i++) {
        seed = seed * 1103515245u + 12345u;
        src[i] = (unsigned char)(seed >> 24);
    }
    const size_t sizes[] = {0, 1, 7, 4096, (1u << 20) - 3, (1u << 20), (3u << 20) + 5, (17u << 20) + 11, cap - 64};
    for (int round = 0; round < 2; round++) {
        for (size_t n : sizes) {
            for (size_t off : {(size_t)0, (size_t)13}) {
                if (off + n > cap) continue;
                std::memset(dst.data(), 0xA5, cap);
                pool.copy(dst.data() + off, src.data() + off, n);
                if (std::memcmp(dst.data() + off, src.data() + off, n) != 0) {
                    std::printf("mismatch: n=%zu off=%zu round=%d\n", n, off, round);
                    return 1;
                }
                // bytes outside [off, off + n) are untouched
                for (size_t i = 0; i < off; i++)
                    if (dst[i] != 0xA5) {
                        std::printf("wrote before the range: n=%zu off=%zu\n", n, off);
                        return 1;
                    }
                if (off + n < cap && dst[off + n] != 0xA5) {
                    std::printf("wrote past the range: n=%zu off=%zu\n", n, off);
                    return 1;
                }
            }
        }
        pool.stop();
        pool.start(round == 0 ? 2 : 0);
    }
    std::printf("ok\n");
    return 0;
}
