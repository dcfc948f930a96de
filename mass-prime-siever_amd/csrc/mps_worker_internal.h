// Helpers shared by the plain (mps_gpu.hip) and batched (mps_batch.hip) device workers.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <atomic>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "mps_gpu.h"
#include "mps_kernels.h"

// One pass lane of a handle: a stream with the buffers, pinned staging and events of one pass in
// flight.  A call that runs passes holds one lane; with two lanes a second feeder thread's pass is
// queued on the device while the first one's records come back and the host works on them, so the
// device does not idle between the feeders' chunks.
constexpr int kPassLanes = 2;
struct PassLane {
    std::mutex mu;  // held by the call that runs passes on this lane
    hipStream_t stream = nullptr;
    mps::PassBuffers pb{};  // pb.cnt is allocated at create, the rest by ensure_buffers
    uint64_t buf_odds = 0;
    uint64_t* d_vals = nullptr;  // the filter path's values
    size_t vals_cap = 0;
    mps::Counters* h_cnt = nullptr;
    uint64_t* h_out = nullptr;
    mps::PassDesc* d_desc = nullptr;
    mps::PassDesc* h_desc = nullptr;
    uint2* d_csz = nullptr;
    uint2* h_csz = nullptr;
    hipEvent_t ev[6] = {};
};

// One device handle: the plain worker's pass lanes, and the device-resident sieving primes that
// the batched worker shares.  Settings (classes, policy, field, pass size, tuning defaults) change
// only while the caller holds every lane.
struct mps_gpu {
    int device = 0;
    bool ok = false;
    bool classes_on = true;
    std::atomic<bool> broken{false};
    int field_mode = -1;
    uint64_t pass_odds = 1ull << 25;
    uint32_t mcap = 2048;
    uint32_t waves = 0;  // 0: by item (4 for plain-mode tables above 1024 baby steps, else 2)
    int class_policy = 1;  // 0: classes for an even cm only (gpu_classes, Math.hpp:324-337); 1: any cm
    uint32_t ext_bound = 0;  // order mode: largest medium prime of p - 1 used (0: by field, 1: none)
    // autotuned table shapes by item key
    mutable std::mutex tune_mu;
    std::map<uint32_t, mps_tuning> tuned;
    PassLane lane[kPassLanes];
    std::atomic<int> lanes{kPassLanes};  // lanes in use: one when two passes' buffers would not fit the device
    size_t mem_bytes = 0;    // the device's memory
    std::atomic<uint32_t> next_lane{0};  // the lane a call waits for when none is free
    mutable std::mutex st_mu;  // st and err, written by calls on different lanes
    std::mutex noalloc_mu;  // one allocation-free call at a time (they share the scratch array)
    mps::SievePrime* d_sp = nullptr;
    mps::SieveTable tab{};  // d_sp with its counts
    // allocation-free entry point: this device's scratch array (device address from
    // hipGetSymbolAddress at create) and a pinned record staging area of kStageRecords records
    void* scratch = nullptr;
    uint64_t* h_stage = nullptr;
    mps_gpu_stats st{};
    std::string err;
    std::string name;
    int num_cu = 0;
};

namespace mps {

// Kernel constants of one item (the conversions of Gpu.hpp:862-875 and gpu_classes,
// Math.hpp:324-337), the largest baby-step count m_max (plain-mode m is capped at mcap) and the
// class-mode size table csz[qp * 32 + s] = (m, giant steps per class).
// Returns false, with the reason, when the BSGS table would not fit.
bool item_params(const mps_item* it, bool classes_on, int class_policy, uint32_t mcap, uint32_t cmax,
                 uint32_t waves, uint32_t ext_bound, ItemParams& ip, uint64_t& m_max, uint2* csz, std::string& why);

// Everything a pass of one item needs besides the range: constants, class-size table, geometry.
struct Plan {
    ItemParams ip{};
    uint64_t m_max = 0;
    TableGeom tg{};
    std::array<uint2, kCszEntries> csz;
};

// Key of the tuning table: class mode and modulus kind, p magnitude (6-bit steps of log2 P0) and
// exponent span per class modulus (2-bit steps of its log2).
uint32_t tune_key(const mps_gpu* g, const mps_item* item, uint64_t P0, bool classes_on);

// The tuning an item gets: the candidate the autotuner is timing on this thread, else the table
// entry of its key, else the handle's settings (waves 0 = by item, mcap, class-mode cap
// kClassMaxM, load 2).
mps_tuning item_tuning(const mps_gpu* g, const mps_item* item, uint64_t P0, bool classes_on);

// Item constants and table choices for an item whose P range starts at P0.
bool plan_item(const mps_gpu* g, const mps_item* item, uint64_t P0, bool classes_on, Plan& pl, std::string& why);

// Field of a pass ending below hi under the handle's setting.
FieldKind pass_field(const mps_gpu* g, uint64_t hi);

}  // namespace mps
