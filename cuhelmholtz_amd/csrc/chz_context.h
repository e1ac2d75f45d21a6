// chz_context.h -- host-side state of the library: per-device twiddle
// tables and axis-pass configurations, error reporting.
#pragma once
#include <hip/hip_runtime.h>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "chz_kernels.h"

namespace chz {

// Launch configuration of one axis pass for one (N, precision, tile shape).
struct AxisCfg {
    int N = 0, M = 0;
    int EPT = 0, TPL = 0, L = 0;
    unsigned long long plan = 0;   // radices, one byte per stage
};

// Kernel roles that keep separate launch configurations.  ROLE_ZMIDN: the
// fused z pass of a solve whose z code is not periodic (its transforms are
// DCT / DST kinds, which want other tile shapes than the periodic pass); a
// size with no configuration of its own for it uses the ROLE_ZMID one.
enum Role { ROLE_ROWS = 0, ROLE_COLS = 1, ROLE_XFWD = 2, ROLE_ZMID = 3, ROLE_ZCOLS = 4, ROLE_ZMIDN = 5, ROLE_COUNT = 6 };

// Status codes of the C ABI (also in include/cuhelmholtz.h).
enum Status { OK = 0, ERR_INVALID = -1, ERR_UNSUPPORTED = -2, ERR_DEVICE = -3, ERR_ALLOC = -4 };

void set_error(const std::string &msg);
int fail(int code, const std::string &msg);
const char *last_error();

class Context {
public:
    static Context &get();           // context of the current device
    // device twiddle table of N at precision prec (64 or 32); nullptr on failure
    const void *twiddles(int N, int prec);
    // launch configuration of a pass kernel (role: see Role)
    bool config(int N, int prec, int role, AxisCfg *out);
    bool config(int N, int prec, bool rows, AxisCfg *out) { return config(N, prec, rows ? ROLE_ROWS : ROLE_COLS, out); }
    // validate and pin a configuration (autotuner); false if it cannot run
    bool set_config(int N, int prec, int role, int ept, int tpl, int lines, unsigned long long *plan_out);
    void reset_configs() { cfg_.clear(); }
    // pin the configurations listed in a tuning table ("role N prec ept tpl lines" per line)
    int load_tuning(const char *path);
    size_t lds_cap() const { return lds_cap_; }
    int device() const { return dev_; }
    // wait for the device, then free the twiddle tables, workspaces and
    // cached tables (the launch configurations are kept)
    void release();
    // persistent workspace `slot` of at least `bytes` (grown on demand, kept
    // until release()), ordered on stream st after the last call that used
    // the slot: st waits for that call's end event (no host wait; growing the
    // slot waits on the host before freeing the old memory).  The caller
    // marks its own end with workspace_done() after its last launch.
    void *workspace(size_t bytes, int slot, hipStream_t st);
    void workspace_done(int slot, hipStream_t st);
    // device copy of a small host table, cached by its bytes (at most
    // kMaxBlobs kept; the least recently used is freed after a device wait)
    void *cached_blob(const std::string &key, const void *data, size_t bytes);
    static const size_t kMaxBlobs = 64;
    std::recursive_mutex &mutex() { return mu_; }
    // per-pass timing (the counterpart of the reference's TIME_TEST splits):
    // when enabled, every pass of a transform / solve records an event at its
    // start and each call one at its end; events accumulate until read.
    void set_timing(bool on) { timing_ = on; nev_ = 0; names_.clear(); }
    bool timing() const { return timing_; }
    void timing_mark(const char *name, hipStream_t s);   // start of a pass named `name`
    void timing_end(hipStream_t s);                      // end of the call's last pass
    // durations of the recorded passes (waits for the events), then clears
    int timing_read(double *ms, int max, std::string *names);

private:
    explicit Context(int dev);
    int dev_;
    size_t lds_cap_ = 65536;
    std::recursive_mutex mu_;   // held while a call enqueues (re-entered by nested calls)
    std::map<std::pair<int, int>, void *> tw_;
    std::map<std::tuple<int, int, int>, AxisCfg> cfg_;
    struct Slot {
        void *p = nullptr;
        size_t bytes = 0;
        hipEvent_t done = nullptr;   // end of the last call that used the slot
        bool recorded = false;
    };
    std::map<int, Slot> ws_;
    std::map<std::string, std::pair<void *, unsigned long long>> blobs_;   // device copy, last use
    unsigned long long blob_clock_ = 0;
    bool timing_ = false;
    std::vector<hipEvent_t> ev_;
    int nev_ = 0;
    std::vector<std::string> names_;
};

// The workspace slots one call uses: acquired in order on the call's stream,
// released (end events recorded) when the lease goes out of scope, after the
// call's launches.
struct WsLease {
    Context &ctx;
    hipStream_t s;
    int slots[8];
    int n = 0;
    WsLease(Context &c, hipStream_t st) : ctx(c), s(st) {}
    void *get(size_t bytes, int slot)
    {
        void *p = ctx.workspace(bytes, slot, s);
        if (p && n < 8) slots[n++] = slot;
        return p;
    }
    ~WsLease()
    {
        for (int i = 0; i < n; i++) ctx.workspace_done(slots[i], s);
    }
};

// records pass boundaries while the context's timing is on
struct PassTimer {
    Context &ctx;
    hipStream_t s;
    PassTimer(Context &c, hipStream_t st) : ctx(c), s(st) {}
    void mark(const char *name) { if (ctx.timing()) ctx.timing_mark(name, s); }
    ~PassTimer() { if (ctx.timing()) ctx.timing_end(s); }
};

// host image of the twiddle table of N (layout TwOffsets) at precision prec
std::vector<char> twiddle_bytes(int N, int prec);

// radix plan of an M-point FFT for tpl threads of at most ept points each
bool radix_plan(int M, int ept, int tpl, unsigned long long *plan);

}  // namespace chz
