// Drives the batch engine over the CPU stand-in device (mock_device.cpp) for a ThreadSanitizer
// build: a mix of items on several slots with graphs and without, admission limits, and a device
// error in the middle of a run.  Exits 0 when every item finished with the expected status.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "mps_batch.h"
#include "mps_gpu.h"
#include "mps_host.h"

extern "C" {
mps_gpu* mock_gpu_create(int num_cu);
void mock_gpu_destroy(mps_gpu* g);
void mock_configure(int fail_at, int pass_us);
int mock_table_mismatch();
}

static int run(mps_gpu* g, int graphs, int policy, int max_active, int fail_at) {
    mock_configure(fail_at, 100);
    mps_batch_opts o{};
    o.slots = 4;
    o.pass_odds = 4096;
    o.use_graphs = graphs;
    o.policy = policy;
    o.max_active = max_active;
    mps_batch* b = mps_batch_create(g, &o);
    if (!b) return 1;
    std::vector<int64_t> ids;
    for (int k = 0; k < 24; ++k) {
        const uint64_t P0 = 1000000 + 77777ull * k;
        mps_work_item w{};
        w.b = 2 + k % 5;
        w.c = (k % 2) ? -25 : 111;
        w.P0 = P0;
        w.P1 = P0 + 8192ull * (1 + k % 4) + 17;
        w.N0 = 0;
        w.N1 = 100000;
        w.cls_mod = (k % 3) ? 2 : 1;
        w.cls_res = (k % 3) ? 1 : 0;
        w.P0_original = P0;
        ids.push_back(mps_batch_submit(b, &w, (uint64_t)k));
    }
    const int rc = mps_batch_wait_all(b);
    std::vector<mps_item_done> done(64);
    size_t n = 0;
    mps_batch_poll(b, done.data(), done.size(), &n);
    int bad = 0;
    if (n != ids.size()) bad++;
    for (size_t i = 0; i < n; ++i) {
        if (fail_at == 0 && done[i].status != MPS_OK) bad++;
        if (fail_at != 0 && done[i].status != MPS_OK && done[i].status != MPS_ERROR) bad++;
    }
    if (fail_at != 0 && rc != MPS_ERROR) bad++;
    if (fail_at == 0 && rc != MPS_OK) bad++;
    mps_batch_destroy(b);
    if (mock_table_mismatch()) bad++;
    std::printf("graphs=%d policy=%d max_active=%d fail_at=%d: %zu items, rc=%d, bad=%d\n", graphs, policy, max_active,
                fail_at, n, rc, bad);
    return bad;
}

int main() {
    mps_gpu* g = mock_gpu_create(8);
    int bad = 0;
    bad += run(g, 1, 0, 0, 0);
    bad += run(g, 0, 1, 0, 0);
    bad += run(g, 1, 1, 3, 0);
    bad += run(g, 1, 0, 0, 7);
    bad += run(g, 0, 0, 2, 3);
    mock_gpu_destroy(g);
    std::printf(bad ? "FAILED\n" : "OK\n");
    return bad ? 1 : 0;
}
