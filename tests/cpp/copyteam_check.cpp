// CPU check of the host copy team behind the pinned staging ring: copies of
// sizes below and above the split threshold, not multiples of the threads'
// share or of 64 bytes, at unaligned offsets, with the team stopped and
// restarted between rounds (a restarted team must not replay the previous
// team's last job).  Every byte of the destination is checked, including
// the guard bytes around the copied range.  Built with ThreadSanitizer.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../natural_number_amd/csrc/nnh_copyteam.hpp"

int main() {
    const uint64_t cap = 24ull << 20;
    std::vector<uint8_t> src(cap + 64), dst(cap + 64);
    uint64_t x = 0x9e3779b97f4a7c15ull;
    for (auto& b : src) {
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        b = (uint8_t)x;
    }
    const uint64_t sizes[] = {0, 1, 63, 4095, (2ull << 20) - 1, 2ull << 20, (2ull << 20) + 1,
                              (16ull << 20) + 7, 23333333, cap};
    nnh::CopyTeam<3> team;
    int bad = 0;
    for (int round = 0; round < 3; round++) {
        for (uint64_t n : sizes)
            for (uint64_t off : {0ull, 3ull})
                for (uint64_t soff : {0ull, 5ull}) {
                    if (off + n > dst.size() || soff + n > src.size()) continue;
                    std::memset(dst.data(), 0xAB, dst.size());
                    team.copy(dst.data() + off, src.data() + soff, n);
                    if (std::memcmp(dst.data() + off, src.data() + soff, n) != 0) bad++;
                    for (uint64_t i = 0; i < off; i++) bad += dst[i] != 0xAB;
                    for (uint64_t i = off + n; i < dst.size(); i++)
                        if (dst[i] != 0xAB) {
                            bad++;
                            break;
                        }
                }
        team.stop();  // the next round starts a new team
    }
    std::printf("%s (%d mismatches)\n", bad ? "FAIL" : "ok", bad);
    return bad ? 1 : 0;
}
