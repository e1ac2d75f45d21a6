// chz_batch.cpp -- many independent Helmholtz solves (uneven sizes, mixed
// boundary codes) replayed as one hipGraph.
//
// A plan is built once from the problem list.  Every problem's solve is the
// same pass sequence as the single solve (5 passes for FP64; 5 per FP32
// solve / refinement for the mixed form), so step s of all problems is
// launched together: one persistent launch per size class (problems sharing
// the cell count -- hence the launch configuration -- along the pass axis).
// Each launch walks a work list of (problem, tile) items and reads every
// problem's pass arguments from a device descriptor table.  The periodic
// copies follow as three launches (x, y, z faces, in that order, as in the
// single solve).  The whole sequence is captured into one hipGraph at plan
// creation; a solve is one hipGraphLaunch.  The node arrays f of the problems
// are bound at creation: a solve reads the right-hand sides from them and
// leaves the solutions there.
#include <algorithm>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "chz_context.h"
#include "chz_solver.h"
#include "chz_transform.h"
#include "../../include/cuhelmholtz.h"

namespace chz {

struct BatchLaunch {
    const void *fn;
    int threads, grid;
    size_t lds;
    const BatchItem *items;   // device
    int nitems;
    const BatchDesc *descs;   // device
    std::string name;
};

struct BatchPlan {
    int dev = 0, count = 0, prec = 64, refine = 0;
    void *work = nullptr;     // per-problem workspaces
    void *tables = nullptr;   // descriptors, items, periodic copies
    size_t work_bytes = 0, table_bytes = 0;
    std::vector<BatchLaunch> launches;
    const BatchCopy *copies[3] = {nullptr, nullptr, nullptr};
    int ncopies[3] = {0, 0, 0}, copy_grid[3] = {0, 0, 0};
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~BatchPlan()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (cap) (void)hipStreamDestroy(cap);
        if (tables) (void)hipFree(tables);
        if (work) (void)hipFree(work);
    }
};

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

static hipError_t launch_batch(const BatchLaunch &l, hipStream_t st)
{
    const BatchItem *items = l.items;
    int n = l.nitems;
    const BatchDesc *descs = l.descs;
    void *args[] = {&items, &n, &descs};
    return hipLaunchKernel(l.fn, dim3((unsigned)l.grid), dim3((unsigned)l.threads), args, l.lds, st);
}

static hipError_t launch_copies(const BatchPlan &p, int axis, hipStream_t st)
{
    if (!p.ncopies[axis]) return hipSuccess;
    const BatchCopy *tab = p.copies[axis];
    void *args[] = {&tab};
    return hipLaunchKernel(batch_periodic_kernel(), dim3((unsigned)p.copy_grid[axis], (unsigned)p.ncopies[axis]),
                           dim3(256), args, 0, st);
}

// enqueue the whole batched solve on st (eager launches; optional pass timing)
static int enqueue(Context &ctx, const BatchPlan &p, hipStream_t st, bool timed)
{
    hipError_t e;
    for (const BatchLaunch &l : p.launches) {
        if (timed) ctx.timing_mark(l.name.c_str(), st);
        if ((e = launch_batch(l, st)) != hipSuccess)
            return fail(ERR_DEVICE, "batched pass " + l.name + " launch: " + hipGetErrorString(e));
    }
    for (int a = 0; a < 3; a++) {
        if (timed && p.ncopies[a]) ctx.timing_mark("periodic", st);
        if ((e = launch_copies(p, a, st)) != hipSuccess)
            return fail(ERR_DEVICE, std::string("batched periodic copy launch: ") + hipGetErrorString(e));
    }
    return OK;
}

static int create(const chz_problem *probs, int count, int prec, int refine, BatchPlan **out)
{
    if (!out) return fail(ERR_INVALID, "null plan pointer");
    *out = nullptr;
    if (!probs || count <= 0) return fail(ERR_INVALID, "empty problem list");
    if (prec != 64 && prec != 32) return fail(ERR_INVALID, "precision must be 64 or 32");
    if (prec == 64 && refine != 0) return fail(ERR_INVALID, "the FP64 batched solve takes no refinement steps");
    if (refine < 0 || refine > 16) return fail(ERR_INVALID, "refinement count must be in 0..16");
    Context &ctx = Context::get();
    std::lock_guard<std::mutex> lock(ctx.mutex());

    std::vector<SolveGeom> geo(count);
    const size_t esz = prec == 64 ? 8 : 4;
    std::vector<size_t> ws_off(count), u_off(count);
    size_t total = 0;
    for (int i = 0; i < count; i++) {
        const chz_problem &q = probs[i];
        if (!q.f) return fail(ERR_INVALID, "problem " + std::to_string(i) + ": null solution array");
        int rc = make_geom(geo[i], q.domain, q.N, q.bc, q.lambda, q.faces);
        if (rc) return fail(rc, "problem " + std::to_string(i) + ": " + last_error());
        if (prec == 32 && refine > 0 && geo[i].singular)
            return fail(ERR_UNSUPPORTED, "problem " + std::to_string(i) +
                                             ": singular problems (lambda = 0, periodic/Neumann-Neumann on every axis) "
                                             "are solved in FP64 or alone in mixed precision");
        ws_off[i] = total;
        total += align256(esz * (size_t)ws_row(q.N[0], esz) * (q.N[1] + 2) * (q.N[2] + 2));
        if (prec == 32 && refine > 0) {
            u_off[i] = total;
            total += align256(sizeof(double) * (size_t)(q.N[0] + 1) * (q.N[1] + 1) * (q.N[2] + 1));
        }
    }
    BatchPlan *plan = new BatchPlan();
    plan->count = count;
    plan->prec = prec;
    plan->refine = refine;
    plan->dev = ctx.device();
    plan->work_bytes = total;
    if (hipMalloc(&plan->work, total) != hipSuccess) {
        delete plan;
        return fail(ERR_ALLOC, "batched workspace allocation failed (" + std::to_string(total) + " bytes)");
    }
    char *wbase = reinterpret_cast<char *>(plan->work);

    // pass sequence of every problem
    std::vector<std::vector<PassSpec>> specs(count);
    int rc = OK;
    for (int i = 0; i < count && rc == OK; i++) {
        const SolveGeom &g = geo[i];
        double *f = probs[i].f;
        std::vector<PassSpec> &v = specs[i];
        if (prec == 64) {
            double *ex = reinterpret_cast<double *>(wbase + ws_off[i]);
            if (!(rc = solve_passes<double>(ctx, g, f, ex, nullptr, nullptr, v))) rc = final_pass<double>(ctx, g, ex, nullptr, f, v);
        } else {
            float *ex = reinterpret_cast<float *>(wbase + ws_off[i]);
            double *U = refine > 0 ? reinterpret_cast<double *>(wbase + u_off[i]) : nullptr;
            if (!(rc = solve_passes<float>(ctx, g, f, ex, nullptr, nullptr, v)))
                rc = final_pass<float>(ctx, g, ex, nullptr, refine > 0 ? U : f, v);
            for (int it = 1; it <= refine && rc == OK; it++)
                if (!(rc = solve_passes<float>(ctx, g, f, ex, U, nullptr, v))) rc = final_pass<float>(ctx, g, ex, U, it == refine ? f : U, v);
        }
        if (rc) rc = fail(rc, "problem " + std::to_string(i) + ": " + last_error());
    }
    if (rc) {
        delete plan;
        return rc;
    }
    const size_t nsteps = specs[0].size();
    for (int i = 1; i < count; i++)
        if (specs[i].size() != nsteps) {
            delete plan;
            return fail(ERR_INVALID, "internal: uneven pass sequences");
        }

    // group step s of all problems by launch configuration
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, plan->dev) != hipSuccess || cus <= 0) cus = 256;
    struct Group {
        int step;
        const PassSpec *first;
        std::vector<int> probs;
    };
    std::vector<Group> groups;
    for (size_t s = 0; s < nsteps; s++) {
        std::map<std::tuple<int, int, int, int, unsigned long long>, size_t> idx;
        for (int i = 0; i < count; i++) {
            const PassSpec &q = specs[i][s];
            auto key = std::make_tuple(q.N, q.ept, q.d.p.TPL, q.d.p.L, q.d.p.plan);
            auto it = idx.find(key);
            if (it == idx.end()) {
                idx[key] = groups.size();
                groups.push_back(Group{(int)s, &q, {i}});
            } else {
                groups[it->second].probs.push_back(i);
            }
        }
    }
    // table layout: per group descriptors then items
    std::vector<size_t> d_off(groups.size()), i_off(groups.size());
    std::vector<int> nitems(groups.size());
    size_t tb = 0;
    for (size_t gi = 0; gi < groups.size(); gi++) {
        const Group &gr = groups[gi];
        long long n = 0;
        for (int i : gr.probs) {
            const PassSpec &q = specs[i][gr.step];
            n += (long long)((q.d.p.na + q.d.p.L - 1) / q.d.p.L) * q.nb;
        }
        if (n > (1LL << 30)) {
            delete plan;
            return fail(ERR_INVALID, "batched pass with more than 2^30 tiles");
        }
        nitems[gi] = (int)n;
        d_off[gi] = tb;
        tb += align256(sizeof(BatchDesc) * gr.probs.size());
        i_off[gi] = tb;
        tb += align256(sizeof(BatchItem) * (size_t)n);
    }
    size_t c_off[3];
    std::vector<BatchCopy> copies[3];
    for (int a = 0; a < 3; a++) {
        int maxface = 0;
        for (int i = 0; i < count; i++) {
            const SolveGeom &g = geo[i];
            if (g.bc[a] != 0) continue;
            copies[a].push_back(BatchCopy{probs[i].f, g.N[0], g.N[1], g.N[2], a});
            const int face = a == 0 ? (g.N[1] + 1) * (g.N[2] + 1) : a == 1 ? (g.N[0] + 1) * (g.N[2] + 1)
                                                                             : (g.N[0] + 1) * (g.N[1] + 1);
            maxface = std::max(maxface, face);
        }
        if (copies[a].size() > 65535) {
            delete plan;
            return fail(ERR_INVALID, "more than 65535 periodic problems along one axis");
        }
        plan->ncopies[a] = (int)copies[a].size();
        plan->copy_grid[a] = std::min(64, (maxface + 255) / 256);
        c_off[a] = tb;
        tb += align256(sizeof(BatchCopy) * copies[a].size());
    }
    std::vector<char> host(tb, 0);
    for (size_t gi = 0; gi < groups.size(); gi++) {
        const Group &gr = groups[gi];
        BatchDesc *d = reinterpret_cast<BatchDesc *>(host.data() + d_off[gi]);
        BatchItem *it = reinterpret_cast<BatchItem *>(host.data() + i_off[gi]);
        int w = 0;
        for (size_t k = 0; k < gr.probs.size(); k++) {
            const PassSpec &q = specs[gr.probs[k]][gr.step];
            d[k] = q.d;
            // tile t = a + ntiles * b; listed a-fastest, or b-fastest where the
            // pass reads neighbour planes (the residual: consecutive items then
            // share rows of U in L2)
            const int ntiles = (q.d.p.na + q.d.p.L - 1) / q.d.p.L;
            if (q.d.p.bfast) {
                for (int a = 0; a < ntiles; a++)
                    for (int b = 0; b < q.nb; b++) it[w++] = BatchItem{(int)k, a + ntiles * b};
            } else {
                for (int t = 0; t < ntiles * q.nb; t++) it[w++] = BatchItem{(int)k, t};
            }
        }
    }
    for (int a = 0; a < 3; a++)
        if (!copies[a].empty()) memcpy(host.data() + c_off[a], copies[a].data(), sizeof(BatchCopy) * copies[a].size());
    plan->table_bytes = tb;
    if (hipMalloc(&plan->tables, tb) != hipSuccess) {
        delete plan;
        return fail(ERR_ALLOC, "batched descriptor table allocation failed");
    }
    if (hipMemcpy(plan->tables, host.data(), tb, hipMemcpyHostToDevice) != hipSuccess) {
        delete plan;
        return fail(ERR_DEVICE, "batched descriptor table upload failed");
    }
    char *tbase = reinterpret_cast<char *>(plan->tables);
    for (int a = 0; a < 3; a++) plan->copies[a] = reinterpret_cast<const BatchCopy *>(tbase + c_off[a]);

    for (size_t gi = 0; gi < groups.size(); gi++) {
        const Group &gr = groups[gi];
        const PassSpec &q = *gr.first;
        BatchLaunch l;
        l.lds = axis_pass_lds(q.prec, q.d.p.M, q.d.p.L, q.d.p.TPL);
        // lean tile I/O where the windows of every problem of the group allow it
        const FixedEntry *ln = batch_line_kernel(q.prec, q.op, q.N, q.ept, q.d.p.TPL, q.d.p.L, q.d.p.plan);
        for (size_t k = 0; ln && k < gr.probs.size(); k++)
            if (!line_windows_ok(specs[gr.probs[k]][gr.step].d.p, q.N)) ln = nullptr;
        if (ln) {
            l.fn = ln->fn;
            l.lds += ln->lds_extra;
        } else {
            l.fn = batch_fixed_kernel(q.prec, q.op, q.N, q.ept, q.d.p.TPL, q.d.p.L, q.d.p.plan);
        }
        if (!l.fn) l.fn = q.prec == 64 ? batch_kernel_f64(q.op, q.ept) : batch_kernel_f32(q.op, q.ept);
        if (!l.fn) {
            delete plan;
            return fail(ERR_UNSUPPORTED, "no batched kernel for this pass");
        }
        l.threads = q.d.p.L * q.d.p.TPL;
        int occ = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, l.fn, l.threads, l.lds) != hipSuccess || occ <= 0) occ = 1;
        long long grid = std::min<long long>(nitems[gi], (long long)occ * cus);
        if (grid >= 8) grid -= grid % 8;   // whole XCD rounds (item chunks per XCD)
        l.grid = (int)std::max<long long>(grid, 1);
        l.items = reinterpret_cast<const BatchItem *>(tbase + i_off[gi]);
        l.nitems = nitems[gi];
        l.descs = reinterpret_cast<const BatchDesc *>(tbase + d_off[gi]);
        l.name = std::string(q.name) + "/" + std::to_string(q.N);
        if (l.nitems > 0) plan->launches.push_back(l);
    }

    // capture the sequence into a graph
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&plan->cap, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipStreamBeginCapture(plan->cap, hipStreamCaptureModeThreadLocal)) != hipSuccess) {
        delete plan;
        return fail(ERR_DEVICE, std::string("graph capture start: ") + hipGetErrorString(e));
    }
    rc = enqueue(ctx, *plan, plan->cap, false);
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(plan->cap, &graph);
    plan->graph = graph;
    if (rc) {
        delete plan;
        return rc;
    }
    if (e != hipSuccess || !graph) {
        delete plan;
        return fail(ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(e));
    }
    if ((e = hipGraphInstantiate(&plan->exec, graph, nullptr, nullptr, 0)) != hipSuccess) {
        delete plan;
        return fail(ERR_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(e));
    }
    *out = plan;
    return OK;
}

}  // namespace chz

using chz::BatchPlan;

extern "C" {

int chz_batch_create(const chz_problem *problems, int count, int precision, int refine, void **plan)
{
    BatchPlan *p = nullptr;
    int rc = chz::create(problems, count, precision, refine, &p);
    if (plan) *plan = rc ? nullptr : p;
    return rc;
}

int chz_batch_solve(void *plan, int use_graph, void *stream)
{
    if (!plan) return chz::fail(chz::ERR_INVALID, "null plan");
    BatchPlan *p = reinterpret_cast<BatchPlan *>(plan);
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev != p->dev) return chz::fail(chz::ERR_INVALID, "plan belongs to another device");
    hipStream_t st = (hipStream_t)stream;
    if (use_graph) {
        hipError_t e = hipGraphLaunch(p->exec, st);
        if (e != hipSuccess) return chz::fail(chz::ERR_DEVICE, std::string("graph launch: ") + hipGetErrorString(e));
        return chz::OK;
    }
    chz::Context &ctx = chz::Context::get();
    std::lock_guard<std::mutex> lock(ctx.mutex());
    const bool tm = ctx.timing();
    int rc = chz::enqueue(ctx, *p, st, tm);
    if (tm) ctx.timing_end(st);
    return rc;
}

int chz_batch_info(void *plan, int *launches, long long *work_bytes, long long *table_bytes)
{
    if (!plan) return chz::fail(chz::ERR_INVALID, "null plan");
    const BatchPlan *p = reinterpret_cast<const BatchPlan *>(plan);
    int n = (int)p->launches.size();
    for (int a = 0; a < 3; a++) n += p->ncopies[a] ? 1 : 0;
    if (launches) *launches = n;
    if (work_bytes) *work_bytes = (long long)p->work_bytes;
    if (table_bytes) *table_bytes = (long long)p->table_bytes;
    return chz::OK;
}

int chz_batch_destroy(void *plan)
{
    delete reinterpret_cast<BatchPlan *>(plan);
    return chz::OK;
}

}  // extern "C"
