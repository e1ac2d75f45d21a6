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
    // axis-pass configuration; rows = x-pass row tiles
    bool config(int N, int prec, bool rows, AxisCfg *out);
    size_t lds_cap() const { return lds_cap_; }
    int device() const { return dev_; }
    void release();
    // persistent workspace (grown on demand, kept until release())
    void *workspace(size_t bytes, int slot = 0);
    std::mutex &mutex() { return mu_; }

private:
    explicit Context(int dev);
    int dev_;
    size_t lds_cap_ = 65536;
    std::mutex mu_;
    std::map<std::pair<int, int>, void *> tw_;
    std::map<std::tuple<int, int, bool>, AxisCfg> cfg_;
    std::map<int, std::pair<void *, size_t>> ws_;
};

// radix plan of M with every radix dividing ept; false if none
bool radix_plan(int M, int ept, unsigned long long *plan);

}  // namespace chz
