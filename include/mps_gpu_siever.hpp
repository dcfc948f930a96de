// mps_gpu_siever.hpp -- C++ drop-in for the reference's GpuSiever (Gpu.hpp:1040-1074) over the
// C-ABI of mps_gpu.h.  Same five methods, same argument meaning, same return convention:
// false = this chunk takes the caller's CPU path; a device error disables the GPU sieve for the
// rest of the process.  Works with any WorkItem type whose b, c, N0, N1 convert to int64_t and
// whose cls_mod, cls_res are unsigned (the reference's WorkItem, Types.hpp:14-26).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "mps_gpu.h"

namespace mps {

class GpuSiever {
public:
    GpuSiever() = default;
    GpuSiever(const GpuSiever&) = delete;
    GpuSiever& operator=(const GpuSiever&) = delete;
    ~GpuSiever() { mps_gpu_destroy(h_); }

    // GpuSiever::init (Gpu.hpp:1042-1056): backend "auto" or "hip"; the banner starts with
    // "GPU enabled (HIP)" or "GPU disabled".
    bool init(int device_index, const std::string& backend = "auto") {
        if (backend != "auto" && backend != "hip") {
            banner_ = "GPU disabled: GPU_backend=" + backend + " is not available in this build";
            return false;
        }
        char buf[512] = {};
        h_ = mps_gpu_create(device_index, buf, sizeof(buf));
        banner_ = buf;
        return h_ != nullptr;
    }
    bool ok() const { return mps_gpu_ok(h_) != 0; }
    const std::string& banner() const { return banner_; }
    void set_classes(bool on) { mps_gpu_set_classes(h_, on ? 1 : 0); }

    // GpuSiever::filter (Gpu.hpp:1058-1061): appends the ascending indices of the primes that may
    // divide an in-class b^n + c.
    template <class WorkItemT>
    bool filter(const std::vector<uint64_t>& primes, const WorkItemT& item, std::vector<std::size_t>& todo) {
        const mps_item it = convert(item);
        std::vector<uint64_t> out(primes.size() / 16 + 1024);
        size_t n = 0;
        int st = mps_gpu_filter(h_, primes.data(), primes.size(), &it, out.data(), out.size(), &n);
        for (int k = 0; st == MPS_CAPACITY && k < kRetries; ++k) {
            out.resize(n + n / 8 + 64);
            st = mps_gpu_filter(h_, primes.data(), primes.size(), &it, out.data(), out.size(), &n);
        }
        if (st != MPS_OK) return false;
        for (size_t i = 0; i < n; ++i) todo.push_back((std::size_t)out[i]);
        return true;
    }

    // GpuSiever::sieve_filter (Gpu.hpp:1064-1067): appends the numbers of [lo|1, hi) the device
    // could not rule out (exact relations' primes plus the few it leaves to the host).
    template <class WorkItemT>
    bool sieve_filter(uint64_t lo, uint64_t hi, const WorkItemT& item, std::vector<uint64_t>& flagged) {
        const mps_item it = convert(item);
        std::vector<uint64_t> out(4096);
        size_t n = 0;
        int st = mps_gpu_sieve_filter(h_, lo, hi, &it, out.data(), out.size(), &n);
        for (int k = 0; st == MPS_CAPACITY && k < kRetries; ++k) {
            out.resize(n + n / 8 + 64);
            st = mps_gpu_sieve_filter(h_, lo, hi, &it, out.data(), out.size(), &n);
        }
        if (st != MPS_OK) return false;
        flagged.insert(flagged.end(), out.begin(), out.begin() + (std::ptrdiff_t)n);
        return true;
    }

    mps_gpu* handle() const { return h_; }

private:
    // A second call returns a count at least as large only where the device's entry list was full
    // (the numbers past it go to the host, and which ones depends on the order of slot
    // reservations), so the capacity retry takes a margin and a few rounds.
    static constexpr int kRetries = 4;

    template <class WorkItemT>
    static mps_item convert(const WorkItemT& item) {
        mps_item it;
        it.b = static_cast<int64_t>(item.b);
        it.c = static_cast<int64_t>(item.c);
        it.N0 = static_cast<int64_t>(item.N0);
        it.N1 = static_cast<int64_t>(item.N1);
        it.cls_mod = static_cast<uint64_t>(item.cls_mod);
        it.cls_res = static_cast<uint64_t>(item.cls_res);
        return it;
    }

    mps_gpu* h_ = nullptr;
    std::string banner_;
};

}  // namespace mps
