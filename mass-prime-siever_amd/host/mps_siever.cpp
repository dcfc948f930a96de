// mps_siever: the command-line siever on the MI355X worker.  Same files as the reference's
// Siever.cpp (Siever.cpp:75-212): reads prime.ini (Config.hpp:50-121) and worktodo.txt
// (Config.hpp:124-200) from the working directory, appends "GPU enabled ..." and progress lines
// to siever.log, writes relations.p ("p, M, q" lines, Worker.hpp:294-297) and appends one object
// per item to results.json (Results.hpp:72-140).  Every chunk runs on the GPU sieve path; a chunk
// the device declines stops the run (exit 1), as GPU_strict=1 does in the reference.
#include <algorithm>
#include <cctype>
#include <csignal>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "mps_batch.h"
#include "mps_gpu.h"
#include "mps_host.h"

namespace {

std::string utc_now() {  // format_utc_timestamp (Logger.hpp:16-29)
    const std::time_t t = std::time(nullptr);
    std::tm tm{};
    gmtime_r(&t, &tm);
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%04d-%02d-%02d %02d:%02d:%02d UTC+0", tm.tm_year + 1900, tm.tm_mon + 1, tm.tm_mday,
                  tm.tm_hour, tm.tm_min, tm.tm_sec);
    return buf;
}

FILE* g_log = nullptr;

void log_line(const char* level, const std::string& msg) {
    const std::string line = utc_now() + " [" + level + "] " + msg;
    std::fprintf(stderr, "%s\n", line.c_str());
    if (g_log) {
        std::fprintf(g_log, "%s\n", line.c_str());
        std::fflush(g_log);
    }
}

[[noreturn]] void fatal(const std::string& msg) {
    log_line("FATAL", msg);
    std::exit(1);
}

std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

struct Config {  // PrimeConfig (Config.hpp:16-26)
    int cpu_count = 0;
    int memory_allowed_mb = 10240;
    int use_gpu = 0;
    int gpu_device = 0;
    std::string gpu_backend = "auto";
    int gpu_threads = 2;
    int gpu_sieve = 0;
    int gpu_classes = 1;
    int gpu_strict = 0;
    int gpu_batch = 0;     // this build: 1 = every item through the batch engine (mps_batch.h)
    int gpu_autotune = 0;  // this build: 1 = autotune each kind of item before running it
};

int hw_threads() {
    const int n = (int)std::thread::hardware_concurrency();
    return n > 0 ? n : 1;
}

Config load_config(const char* path) {  // load_prime_config (Config.hpp:50-121)
    Config cfg;
    cfg.cpu_count = hw_threads();
    bool has_cpu = false, has_mem = false;
    std::ifstream in(path);
    const bool existed = (bool)in;
    std::string line;
    while (existed && std::getline(in, line)) {
        line = trim(line);
        if (line.empty() || line[0] == '#') continue;
        const size_t eq = line.find('=');
        if (eq == std::string::npos) continue;
        const std::string k = trim(line.substr(0, eq)), v = trim(line.substr(eq + 1));
        if (k.empty()) continue;
        try {
            if (k == "CPU_count") {
                cfg.cpu_count = std::stoi(v);
                has_cpu = true;
            } else if (k == "UseGPU") {
                cfg.use_gpu = std::stoi(v);
            } else if (k == "GPU_device") {
                cfg.gpu_device = std::stoi(v);
            } else if (k == "GPU_backend") {
                cfg.gpu_backend = v;
            } else if (k == "GPU_threads") {
                cfg.gpu_threads = std::max(1, std::stoi(v));
            } else if (k == "GPU_sieve") {
                cfg.gpu_sieve = std::stoi(v);
            } else if (k == "GPU_classes") {
                cfg.gpu_classes = std::stoi(v);
            } else if (k == "GPU_strict") {
                cfg.gpu_strict = std::stoi(v);
            } else if (k == "GPU_batch") {
                cfg.gpu_batch = std::stoi(v);
            } else if (k == "GPU_autotune") {
                cfg.gpu_autotune = std::stoi(v);
            } else if (k == "MemoryAllowed") {
                cfg.memory_allowed_mb = std::stoi(v);
                has_mem = true;
            }
        } catch (const std::exception&) {
            fatal("prime.ini: bad value for " + k);
        }
    }
    if (!(cfg.cpu_count == 0 && cfg.use_gpu != 0) && cfg.cpu_count <= 0) {
        cfg.cpu_count = hw_threads();
        has_cpu = false;
    }
    if (cfg.memory_allowed_mb <= 0) {
        cfg.memory_allowed_mb = 10240;
        has_mem = false;
    }
    if (!has_cpu || !has_mem || !existed) {
        std::ofstream out(path, std::ios::trunc);
        if (!out) fatal(std::string("Failed to write prime.ini at ") + path);
        out << "CPU_count=" << cfg.cpu_count << "\nMemoryAllowed=" << cfg.memory_allowed_mb << "\nUseGPU=" << cfg.use_gpu
            << "\nGPU_device=" << cfg.gpu_device << "\n";
        log_line("INFO", "Wrote prime.ini with CPU_count=" + std::to_string(cfg.cpu_count) +
                             " MemoryAllowed=" + std::to_string(cfg.memory_allowed_mb) + " MB");
    }
    return cfg;
}

bool parse_i64(const std::string& s, int64_t& v) {
    char* end = nullptr;
    errno = 0;
    const long long x = std::strtoll(s.c_str(), &end, 10);
    if (errno || end == s.c_str() || *end) return false;
    v = x;
    return true;
}

std::vector<mps_work_item> load_worktodo(const char* path) {  // Config.hpp:124-200
    std::ifstream in(path);
    if (!in) fatal(std::string("Failed to open worktodo file: ") + path);
    std::vector<mps_work_item> items;
    std::string line;
    size_t line_no = 0;
    while (std::getline(in, line)) {
        ++line_no;
        line = trim(line);
        if (line.empty() || line[0] == '#') continue;
        if (line.rfind("Factor=", 0) == 0) line = line.substr(7);
        std::vector<std::string> parts;
        std::stringstream ss(line);
        std::string tok;
        while (std::getline(ss, tok, ',')) {
            tok = trim(tok);
            if (!tok.empty()) parts.push_back(tok);
        }
        const std::string where = "worktodo line " + std::to_string(line_no);
        if (parts.size() != 6 && parts.size() != 8)
            fatal(where + " must have 6 or 8 comma-separated values (b,c,P0,P1,N0,N1[,mod,res])");
        int64_t v[8] = {0, 0, 0, 0, 0, 0, 1, 0};
        for (size_t i = 0; i < parts.size(); ++i)
            if (!parse_i64(parts[i], v[i])) fatal(where + " parse error: " + parts[i]);
        mps_work_item it{};
        it.b = v[0];
        it.c = v[1];
        if (v[2] < 0 || v[3] < 0) fatal(where + " requires P0, P1 >= 0");
        it.P0 = (uint64_t)v[2];
        it.P1 = (uint64_t)v[3];
        it.P0_original = it.P0;
        it.N0 = v[4];
        it.N1 = v[5];
        it.cls_mod = 1;
        it.cls_res = 0;
        if (parts.size() == 8) {
            if (v[6] <= 0 || v[7] < 0 || v[7] >= v[6]) fatal(where + " parse error: class must satisfy 0 <= res < mod");
            it.cls_mod = (uint64_t)v[6];
            it.cls_res = (uint64_t)v[7];
        }
        if (it.P0 >= it.P1) fatal(where + " requires P0 < P1");
        if (it.N0 >= it.N1) fatal(where + " requires N0 < N1");
        if (it.N1 - it.N0 > 10000000) fatal(where + " range N1-N0 exceeds 10 million");
        items.push_back(it);
    }
    if (items.empty()) fatal(std::string("No work items found in ") + path);
    return items;
}

void append_results(const std::string& obj) {  // Results.hpp:109-139, as a compact JSON array
    std::string content;
    {
        std::ifstream in("results.json", std::ios::binary);
        if (in) content.assign(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
    }
    while (!content.empty() && std::isspace((unsigned char)content.back())) content.pop_back();
    std::string out;
    if (content.empty() || content == "[]") {
        out = "[" + obj + "]";
    } else {
        if (content.front() != '[' || content.back() != ']') fatal("results.json must contain a JSON array");
        content.pop_back();
        out = content + "," + obj + "]";
    }
    std::ofstream f("results.json", std::ios::trunc | std::ios::binary);
    if (!f) fatal("Failed to write results.json");
    f << out << '\n';
}

// ---- checkpointing through per-feeder core files (Worker.hpp:294-297, :369-463)

// Raised by SIGINT / SIGTERM / SIGHUP (Siever.cpp:67-90): no further chunk starts, the finished
// chunks stay in core_<id>.p and the next run resumes from them.
volatile sig_atomic_t g_stop = 0;
static_assert(sizeof(sig_atomic_t) == sizeof(int), "the stop flag is read as an int");

extern "C" void on_shutdown_signal(int) { g_stop = 1; }

std::string core_path(int id) { return "core_" + std::to_string(id) + ".p"; }

// Core files in id order; like the reference's scan, a missing id ends the scan only when the
// next id is missing too.
std::vector<int> core_ids() {
    std::vector<int> ids;
    for (int id = 0;; ++id) {
        if (std::ifstream(core_path(id))) {
            ids.push_back(id);
        } else if (!std::ifstream(core_path(id + 1))) {
            break;
        }
    }
    return ids;
}

bool parse_relation(const std::string& line, mps_relation& r) {
    unsigned long long p = 0, M = 0, q = 0;
    if (std::sscanf(line.c_str(), "%llu, %llu, %llu", &p, &M, &q) != 3) return false;
    r.p = p;
    r.M = M;
    r.q = q;
    return true;
}

// read_checkpoint: the resume point is the smallest last-line p over the core files; lines at or
// above it are dropped (their chunks run again).  Returns the relations kept, or 0 when there is no
// checkpoint ahead of P0 (the item then starts from P0).
uint64_t read_checkpoint(mps_work_item& it) {
    uint64_t new_p0 = it.P1;
    bool any = false;
    const std::vector<int> ids = core_ids();
    for (int id : ids) {
        std::ifstream in(core_path(id));
        std::string line, last;
        while (std::getline(in, line))
            if (!line.empty()) last = line;
        mps_relation r;
        if (!last.empty() && parse_relation(last, r) && r.p < new_p0) {
            new_p0 = r.p;
            any = true;
        }
    }
    if (!any) {
        log_line("INFO", "No checkpoint found, starting from original P0.");
        return 0;
    }
    if (new_p0 <= it.P0) {
        log_line("INFO", "Checkpoint not ahead of current P0, starting from original P0.");
        return 0;
    }
    uint64_t kept_total = 0;
    for (int id : ids) {
        std::vector<std::string> kept;
        {
            std::ifstream in(core_path(id));
            std::string line;
            mps_relation r;
            while (std::getline(in, line))
                if (!line.empty() && parse_relation(line, r) && r.p < new_p0) kept.push_back(line);
        }
        std::ofstream out(core_path(id), std::ios::trunc);
        if (!out) fatal("Failed to trim checkpoint file: " + core_path(id));
        for (const std::string& l : kept) out << l << '\n';
        kept_total += kept.size();
    }
    it.P0 = new_p0;
    log_line("INFO", "Checkpoint recovered: new P0 = " + std::to_string(new_p0) + ", relations restored = " +
                         std::to_string(kept_total));
    return kept_total;
}

struct CoreSink {
    int first_id;  // core_<first_id + feeder>.p, as core_<cpu_count + g>.p (Worker.hpp:340)
    std::mutex mu[64];
};

// Appends one finished chunk's relations to its feeder's core file.
extern "C" void append_core(void* ctx, int feeder, uint64_t, uint64_t, const mps_relation* rels, size_t n) {
    if (n == 0) return;
    CoreSink* sink = static_cast<CoreSink*>(ctx);
    std::lock_guard<std::mutex> l(sink->mu[feeder & 63]);
    FILE* f = std::fopen(core_path(sink->first_id + feeder).c_str(), "ab");
    if (!f) fatal("Failed to append to " + core_path(sink->first_id + feeder));
    for (size_t i = 0; i < n; ++i)
        std::fprintf(f, "%llu, %llu, %llu\n", (unsigned long long)rels[i].p, (unsigned long long)rels[i].M,
                     (unsigned long long)rels[i].q);
    std::fclose(f);
}

// merge_core_files + load_relations (Worker.hpp:465-515): every core file's relations, sorted by p.
std::vector<mps_relation> load_core_relations() {
    std::vector<mps_relation> rels;
    for (int id : core_ids()) {
        std::ifstream in(core_path(id));
        std::string line;
        mps_relation r;
        while (std::getline(in, line))
            if (!line.empty() && parse_relation(line, r)) rels.push_back(r);
    }
    std::sort(rels.begin(), rels.end(), [](const mps_relation& a, const mps_relation& b) { return a.p < b.p; });
    return rels;
}

void remove_core_files() {
    for (int id : core_ids()) std::remove(core_path(id).c_str());
}

// relations.p, the results.json object and the log line of one finished item.
void write_item(const mps_work_item& it, size_t w, size_t nitems, const std::vector<mps_relation>& rels,
                const std::string& stats) {
    FILE* f = std::fopen("relations.p", "wb");
    if (!f) fatal("Failed to create relations.p");
    for (const mps_relation& r : rels)
        std::fprintf(f, "%llu, %llu, %llu\n", (unsigned long long)r.p, (unsigned long long)r.M, (unsigned long long)r.q);
    std::fclose(f);
    log_line("INFO", "Loaded " + std::to_string(rels.size()) + " relation(s)");
    const std::string ts = utc_now();
    const size_t len = mps_results_json(&it, rels.data(), rels.size(), ts.c_str(), nullptr, 0);
    std::string obj(len + 1, '\0');
    mps_results_json(&it, rels.data(), rels.size(), ts.c_str(), &obj[0], obj.size());
    obj.resize(len);
    append_results(obj);
    log_line("INFO", "Finished work item " + std::to_string(w + 1) + "/" + std::to_string(nitems) + ": " +
                         std::to_string(rels.size()) + " relations, " + stats);
}

// Every item through the batch engine at once; outputs are still written in worktodo order.
void run_batched(mps_gpu* gpu, const std::vector<mps_work_item>& items) {
    mps_batch_opts o{};
    o.use_graphs = 1;
    mps_batch* b = mps_batch_create(gpu, &o);
    if (b == nullptr) fatal(std::string("batch engine: ") + mps_gpu_last_error(gpu));
    std::vector<int64_t> ids;
    for (size_t w = 0; w < items.size(); ++w) ids.push_back(mps_batch_submit(b, &items[w], w));
    log_line("INFO", "Submitted " + std::to_string(items.size()) + " work item(s) to the batch engine");
    std::map<uint64_t, mps_item_done> done;
    std::vector<mps_item_done> buf(256);
    for (size_t w = 0; w < items.size(); ++w) {
        while (done.find((uint64_t)ids[w]) == done.end()) {
            size_t n = 0;
            mps_batch_poll(b, buf.data(), buf.size(), &n);
            for (size_t k = 0; k < n; ++k) done[buf[k].id] = buf[k];
            if (n == 0) std::this_thread::sleep_for(std::chrono::milliseconds(1));
        }
        const mps_item_done& d = done[(uint64_t)ids[w]];
        if (d.status != MPS_OK)
            fatal(std::string("GPU worker failed on work item ") + std::to_string(w + 1) + ": status " +
                  std::to_string(d.status) + " " + mps_batch_last_error(b));
        std::vector<mps_relation> rels(d.relations + 1);
        size_t n = 0;
        if (mps_batch_relations(b, d.id, rels.data(), rels.size(), &n) != MPS_OK) fatal("batch relations read failed");
        rels.resize(n);
        char msg[160];
        std::snprintf(msg, sizeof(msg), "submitted at %.3f s, admitted at %.3f s, done at %.3f s", d.t_submit, d.t_admit,
                      d.t_done);
        write_item(items[w], w, items.size(), rels, msg);
    }
    mps_batch_destroy(b);
}

}  // namespace

int main() {
    {
        struct sigaction action {};
        action.sa_handler = on_shutdown_signal;
        sigemptyset(&action.sa_mask);
        sigaction(SIGINT, &action, nullptr);
        sigaction(SIGTERM, &action, nullptr);
        sigaction(SIGHUP, &action, nullptr);
    }
    const auto t0 = std::chrono::steady_clock::now();
    g_log = std::fopen("siever.log", "a");
    log_line("INFO", "Mass Prime Siever (MI355X worker) starting");
    const Config cfg = load_config("prime.ini");
    const std::vector<mps_work_item> items = load_worktodo("worktodo.txt");
    log_line("INFO", "Loaded " + std::to_string(items.size()) + " work item(s)");

    mps_gpu* gpu = nullptr;
    if (cfg.use_gpu != 0) {
        char banner[512] = {};
        if (cfg.gpu_backend == "auto" || cfg.gpu_backend == "hip") gpu = mps_gpu_create(cfg.gpu_device, banner, sizeof(banner));
        else std::snprintf(banner, sizeof(banner), "GPU disabled: GPU_backend=%s is not available in this build", cfg.gpu_backend.c_str());
        log_line("INFO", banner);
    }
    if (gpu == nullptr) fatal("no usable GPU: this siever has no CPU sieve path (set UseGPU=1 on an MI355X)");
    mps_gpu_set_classes(gpu, cfg.gpu_classes != 0);
    if (cfg.gpu_sieve == 0) log_line("INFO", "GPU_sieve=0 requested: this build sieves on the GPU (GPU_sieve=1 behaviour)");
    log_line("INFO", "GPU_threads=" + std::to_string(cfg.gpu_threads) + ", GPU_sieve=1, GPU_classes=" +
                         std::to_string(cfg.gpu_classes) + ", GPU_strict=" + std::to_string(cfg.gpu_strict));
    // worker_chunk_size_bytes (Worker.hpp:152-157)
    const unsigned long long total = (unsigned long long)cfg.memory_allowed_mb << 20;
    const unsigned long long chunk = std::max(1ULL, total / (unsigned long long)std::max(1, cfg.cpu_count));

    std::remove("relations.p");
    if (cfg.gpu_autotune) {
        for (const mps_work_item& it : items) {
            const mps_item di{it.b, it.c, it.N0, it.N1, it.cls_mod, it.cls_res};
            mps_tuning t{};
            double best = 0, dflt = 0;
            if (mps_gpu_autotune(gpu, &di, it.P0, 1ull << 24, &t, &best, &dflt) == MPS_OK) {
                char msg[200];
                std::snprintf(msg, sizeof(msg), "Autotuned b=%lld c=%lld: class policy %u, waves %u, mcap %u, cmax %u, load %u (%.3f ms, default %.3f ms)",
                              (long long)it.b, (long long)it.c, t.policy, t.waves, t.mcap, t.cmax, t.load, best, dflt);
                log_line("INFO", msg);
            }
        }
    }
    if (cfg.gpu_batch) {
        run_batched(gpu, items);
    }
    CoreSink sink;
    sink.first_id = std::max(0, cfg.cpu_count);
    bool stopped = false;
    for (size_t w = 0; w < items.size() && !cfg.gpu_batch; ++w) {
        if (g_stop) {
            log_line("INFO", "Shutdown requested, stopping before next work item.");
            stopped = true;
            break;
        }
        const mps_work_item& it = items[w];
        log_line("INFO", "Processing work item " + std::to_string(w + 1) + "/" + std::to_string(items.size()) +
                             " b=" + std::to_string(it.b) + " c=" + std::to_string(it.c));
        std::remove("relations.p");
        mps_work_item run = it;  // resumes from the checkpoint, if any
        read_checkpoint(run);
        mps_run_opts opts{chunk, cfg.gpu_threads, append_core, &sink, (const volatile int*)&g_stop};
        mps_run_stats st{};
        size_t n = 0;
        const int rc = mps_run_item(gpu, &run, &opts, nullptr, 0, &n, &st);
        if (rc == MPS_STOPPED) {
            log_line("INFO", "Shutdown: checkpoint preserved in core_*.p files");
            stopped = true;
            break;
        }
        if (rc != MPS_OK && rc != MPS_CAPACITY) fatal(std::string("GPU worker failed on this item: ") + mps_gpu_last_error(gpu));
        const std::vector<mps_relation> rels = load_core_relations();
        char msg[160];
        std::snprintf(msg, sizeof(msg), "%llu chunks, %.3f s (device calls %.3f s)", (unsigned long long)st.chunks,
                      st.wall_s, st.device_call_s);
        write_item(it, w, items.size(), rels, msg);
        remove_core_files();
    }
    const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (!stopped) log_line("INFO", "Siever completed successfully, total elapsed " + std::to_string((long long)el) + " s");
    mps_gpu_destroy(gpu);
    if (g_log) std::fclose(g_log);
    return 0;
}
