// chz_context.cpp -- twiddle tables, axis configurations, workspaces, errors.
#include "chz_context.h"

#include <dlfcn.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <tuple>

namespace chz {

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }
int fail(int code, const std::string &msg)
{
    g_err = msg;
    // a failed HIP call (out of memory, launch error) also leaves HIP's
    // per-thread last error set: clear it, the status carries the failure
    // (otherwise the caller's next HIP user, e.g. torch, reports it again)
    if (code == ERR_ALLOC || code == ERR_DEVICE) (void)hipGetLastError();
    return code;
}
const char *last_error() { return g_err.c_str(); }

// Radix sequence for an M-point FFT with TPL threads per line and at most ept
// points per thread: every stage R with a specialised butterfly must satisfy
// ceil(M/(R*TPL)) <= ept/R.  Factors without one (primes from 7 up) become
// direct stages (fft_stage_any: byte R for R <= 255, byte 1 for a remaining
// factor whose smallest prime is larger), which need only M <= ept * TPL.
bool radix_plan(int M, int ept, int tpl, unsigned long long *plan)
{
    static const int radices[] = {16, 12, 10, 8, 6, 5, 4, 3, 2};
    if (M < 1 || M > ept * tpl) return false;
    unsigned long long p = 0;
    int rem = M, st = 0;
    while (rem > 1) {
        int pick = 0;
        for (int R : radices) {
            if (R > ept || rem % R) continue;
            const int nb = (M / R + tpl - 1) / tpl;
            if (nb > ept / R) continue;
            pick = R;
            break;
        }
        if (!pick) {
            int q = 2;
            while (q * q <= rem && rem % q) q++;
            if (rem % q) q = rem;   // rem is prime
            if (q < 7) return false;   // a 2, 3 or 5 that does not fit this thread shape
            pick = q <= 255 ? q : rem;
        }
        if (st >= 8) return false;
        p |= (unsigned long long)(pick <= 255 ? pick : 1) << (8 * st++);
        rem /= pick;
    }
    *plan = p;
    return true;
}

int Context::load_tuning(const char *path)
{
    std::ifstream in(path);
    if (!in) return -1;
    std::string line;
    int n = 0;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream ss(line);
        int role, N, prec, ept, tpl, lines;
        if (!(ss >> role >> N >> prec >> ept >> tpl >> lines)) continue;
        if (set_config(N, prec, role, ept, tpl, lines, nullptr)) n++;
    }
    return n;
}

// directory of this shared library (for the tuning tables shipped beside it)
static std::string library_dir()
{
    Dl_info info;
    if (dladdr((void *)&library_dir, &info) && info.dli_fname) {
        std::string p(info.dli_fname);
        size_t k = p.rfind('/');
        if (k != std::string::npos) return p.substr(0, k);
    }
    return ".";
}

Context::Context(int dev) : dev_(dev)
{
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
        lds_cap_ = (size_t)v;
    if (const char *e = getenv("CHZ_LDS_CAP")) {
        size_t c = (size_t)atol(e);
        if (c > 0 && c < lds_cap_) lds_cap_ = c;
    }
    // tuned launch configurations: CHZ_TUNE_FILE, else tuning/<arch>.txt beside the library
    std::string path;
    if (const char *e = getenv("CHZ_TUNE_FILE")) {
        path = e;
    } else {
        hipDeviceProp_t prop;
        std::string arch = "gfx950";
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess) {
            arch = prop.gcnArchName;
            size_t k = arch.find(':');
            if (k != std::string::npos) arch = arch.substr(0, k);
        }
        path = library_dir() + "/tuning/" + arch + ".txt";
    }
    if (!getenv("CHZ_NO_TUNING")) load_tuning(path.c_str());
}

Context &Context::get()
{
    static std::mutex m;
    static std::map<int, Context *> all;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    std::lock_guard<std::mutex> g(m);
    auto it = all.find(dev);
    if (it != all.end()) return *it->second;
    Context *c = new Context(dev);
    all[dev] = c;
    return *c;
}

// Twiddles are generated in FP64 with the host libm and rounded once to the
// pass precision (the FP32 path uses FP64-generated, rounded twiddles).
std::vector<char> twiddle_bytes(int N, int prec)
{
    const int M = N / 2;
    const int n = TwOffsets::total(N);
    std::vector<double> h(2 * (size_t)n, 0.0);
    const double pi = 3.14159265358979323846;
    for (int k = 0; k <= 2 * N; k++) {
        double a = pi * (double)k / (2.0 * N);
        h[2 * k] = std::cos(a);
        h[2 * k + 1] = std::sin(a);
    }
    for (int j = 0; j < M; j++) {
        double a = 2.0 * pi * (double)j / (double)M;
        h[2 * (TwOffsets::fftw(N) + j)] = std::cos(a);
        h[2 * (TwOffsets::fftw(N) + j) + 1] = -std::sin(a);
    }
    for (int k = 0; k <= M; k++) {
        double a = pi * (double)k / (double)N;
        h[2 * (TwOffsets::rsin(N) + k)] = k == 0 ? 0.0 : 1.0 / std::sin(a);
    }
    std::vector<char> out(h.size() * (prec == 64 ? 8 : 4));
    if (prec == 64) {
        memcpy(out.data(), h.data(), out.size());
    } else {
        float *f = reinterpret_cast<float *>(out.data());
        for (size_t i = 0; i < h.size(); i++) f[i] = (float)h[i];
    }
    return out;
}

const void *Context::twiddles(int N, int prec)
{
    auto key = std::make_pair(N, prec);
    auto it = tw_.find(key);
    if (it != tw_.end()) return it->second;
    const std::vector<char> h = twiddle_bytes(N, prec);
    void *d = nullptr;
    if (hipMalloc(&d, h.size()) != hipSuccess) return nullptr;
    if (hipMemcpy(d, h.data(), h.size(), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(d);
        return nullptr;
    }
    tw_[key] = d;
    return d;
}

bool Context::set_config(int N, int prec, int role, int ept, int tpl, int lines, unsigned long long *plan_out)
{
    static const int epts[] = {16, 12, 10, 8, 6, 5, 4};
    bool ok = false;
    for (int e : epts) ok |= e == ept;
    const int M = N / 2;
    if (!ok || N < 2 || (N & 1) || tpl < 1 || lines < 1 || tpl * lines > CHZ_MAX_THREADS || M > ept * tpl) return false;
    if (role < 0 || role >= ROLE_COUNT || (prec != 64 && prec != 32)) return false;
    unsigned long long p;
    if (!radix_plan(M, ept, tpl, &p)) return false;
    if (axis_pass_lds(prec, M, lines, tpl) > lds_cap_) return false;
    AxisCfg c;
    c.N = N; c.M = M; c.EPT = ept; c.TPL = tpl; c.L = lines; c.plan = p;
    cfg_[std::make_tuple(N, prec, role)] = c;
    if (plan_out) *plan_out = p;
    return true;
}

bool Context::config(int N, int prec, int role, AxisCfg *out)
{
    auto key = std::make_tuple(N, prec, role);
    auto it = cfg_.find(key);
    if (it != cfg_.end()) {
        *out = it->second;
        return true;
    }
    // the non-periodic fused z pass without a configuration of its own: the
    // periodic one's (not cached, so a later change of that one is followed)
    if (role == ROLE_ZMIDN) return config(N, prec, ROLE_ZMID, out);
    const int M = N / 2;
    // 8 points per thread first (measured best for FP64 on gfx950), then larger / smaller
    static const int epts[] = {8, 16, 12, 10, 6, 5, 4};
    AxisCfg c;
    c.N = N;
    c.M = M;
    for (int e : epts) {
        if (c.EPT) break;
        for (int tpl = (M + e - 1) / e; tpl <= CHZ_MAX_THREADS && tpl <= (M > 1 ? M : 1); tpl++) {
            unsigned long long p;
            if (!radix_plan(M, e, tpl, &p)) continue;
            c.EPT = e;
            c.TPL = tpl;
            c.plan = p;
            break;
        }
    }
    if (!c.EPT) return false;
    int L = 256 / c.TPL;
    if (L < 1) L = 1;
    // two workgroups per CU: half of the per-CU LDS (160 KB on gfx950), at most 80 KB
    const size_t cap = lds_cap_ / 2 < 81920 ? (lds_cap_ / 2 > 32768 ? lds_cap_ / 2 : 32768) : 81920;
    while (L > 1 && axis_pass_lds(prec, M, L, c.TPL) > cap) L--;
    if (axis_pass_lds(prec, M, L, c.TPL) > lds_cap_) return false;
    c.L = L;
    cfg_[key] = c;
    *out = c;
    return true;
}

void *Context::workspace(size_t bytes, int slot, hipStream_t st)
{
    Slot &w = ws_[slot];
    if (!w.done && hipEventCreateWithFlags(&w.done, hipEventDisableTiming) != hipSuccess) {
        w.done = nullptr;
        return nullptr;
    }
    if (w.p && w.bytes >= bytes) {
        // stream order after the slot's last user (a no-op on the same stream)
        if (w.recorded && hipStreamWaitEvent(st, w.done, 0) != hipSuccess) return nullptr;
        return w.p;
    }
    if (w.p) {
        // the last user may still run on another stream
        if (w.recorded) (void)hipEventSynchronize(w.done);
        (void)hipFree(w.p);
        w.p = nullptr;
        w.bytes = 0;
    }
    w.recorded = false;
    void *d = nullptr;
    if (hipMalloc(&d, bytes) != hipSuccess) return nullptr;
    w.p = d;
    w.bytes = bytes;
    return d;
}

void Context::workspace_done(int slot, hipStream_t st)
{
    auto it = ws_.find(slot);
    if (it == ws_.end() || !it->second.done) return;
    it->second.recorded = hipEventRecord(it->second.done, st) == hipSuccess;
}

void *Context::cached_blob(const std::string &key, const void *data, size_t bytes)
{
    auto it = blobs_.find(key);
    if (it != blobs_.end()) {
        it->second.second = ++blob_clock_;
        return it->second.first;
    }
    if (blobs_.size() >= kMaxBlobs) {
        // drop the least recently used table once no queued work can read it
        auto lru = blobs_.begin();
        for (auto jt = blobs_.begin(); jt != blobs_.end(); ++jt)
            if (jt->second.second < lru->second.second) lru = jt;
        (void)hipDeviceSynchronize();
        (void)hipFree(lru->second.first);
        blobs_.erase(lru);
    }
    void *d = nullptr;
    if (hipMalloc(&d, bytes) != hipSuccess) return nullptr;
    if (hipMemcpy(d, data, bytes, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(d);
        return nullptr;
    }
    blobs_[key] = std::make_pair(d, ++blob_clock_);
    return d;
}

void Context::timing_mark(const char *name, hipStream_t s)
{
    if ((int)ev_.size() <= nev_) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return;
        ev_.push_back(e);
    }
    if (hipEventRecord(ev_[nev_], s) != hipSuccess) return;
    nev_++;
    names_.push_back(name);
}

void Context::timing_end(hipStream_t s) { timing_mark("", s); }

int Context::timing_read(double *ms, int max, std::string *names)
{
    if (nev_ == 0) return 0;
    if (hipEventSynchronize(ev_[nev_ - 1]) != hipSuccess) return -1;
    int k = 0;
    for (int i = 0; i + 1 < nev_ && k < max; i++) {
        if (names_[i].empty()) continue;           // end-of-call marker
        float t = 0;
        if (hipEventElapsedTime(&t, ev_[i], ev_[i + 1]) != hipSuccess) return -1;
        ms[k] = t;
        if (names) names[k] = names_[i];
        k++;
    }
    nev_ = 0;
    names_.clear();
    return k;
}

void Context::release()
{
    std::lock_guard<std::recursive_mutex> lock(mu_);
    (void)hipDeviceSynchronize();
    for (auto &kv : blobs_) (void)hipFree(kv.second.first);
    blobs_.clear();
    for (auto &kv : tw_) (void)hipFree(kv.second);
    tw_.clear();
    for (auto &kv : ws_) {
        if (kv.second.p) (void)hipFree(kv.second.p);
        kv.second.p = nullptr;
        kv.second.bytes = 0;
        kv.second.recorded = false;
    }
}

}  // namespace chz
