// finish_records_parallel against finish_records on the same device records: every odd number of
// a range is a record the host decides (as the CPU stand-in device emits them), some numbers twice,
// and the relations, flagged-prime counts and status must be identical for 1 to 6 threads.
#include <cstdio>
#include <cstring>
#include <vector>

#include "mps_gpu.h"
#include "mps_host.h"
#include "../mass-prime-siever_amd/host/mps_host_internal.h"

int main() {
    mps_work_item it;
    std::memset(&it, 0, sizeof(it));
    it.b = 2;
    it.c = -25;
    it.P0 = 1000001;
    it.P1 = 1000001 + (1u << 18);
    it.N0 = 3600000;
    it.N1 = 5000000;
    it.cls_mod = 2;
    it.cls_res = 1;
    std::vector<mps_dev_relation> base;
    for (uint64_t x = it.P0 | 1; x < it.P1; x += 2) {
        base.push_back(mps_dev_relation{x, MPS_HOST_DECIDES});
        if (x % 7 == 1) base.push_back(mps_dev_relation{x, MPS_HOST_DECIDES});  // the same number twice
    }
    std::vector<mps_dev_relation> a = base;
    std::vector<mps_relation> want;
    uint64_t want_flagged = 0;
    const int want_st = mps::finish_records(&it, it.P0, it.P1, a.data(), a.size(), want, &want_flagged);
    if (want.empty()) {
        std::printf("no relations in the test range\n");
        return 1;
    }
    for (int threads = 1; threads <= 6; ++threads) {
        std::vector<mps_dev_relation> b(base.rbegin(), base.rend());  // another input order
        std::vector<mps_relation> got;
        uint64_t flagged = 0;
        const int st = mps::finish_records_parallel(&it, it.P0, it.P1, b.data(), b.size(), got, &flagged, threads, 512);
        bool same = st == want_st && flagged == want_flagged && got.size() == want.size();
        for (size_t i = 0; same && i < got.size(); ++i)
            same = got[i].p == want[i].p && got[i].M == want[i].M && got[i].q == want[i].q;
        std::printf("threads %d: %zu relations, %llu flagged, status %d: %s\n", threads, got.size(),
                    (unsigned long long)flagged, st, same ? "same" : "DIFFERENT");
        if (!same) return 1;
    }
    std::printf("ok %zu relations\n", want.size());
    return 0;
}
