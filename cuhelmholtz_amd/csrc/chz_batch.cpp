// chz_batch.cpp -- many independent Helmholtz solves (uneven sizes, mixed
// boundary codes) replayed as one hipGraph.
//
// A plan is built once from the problem list.  Every problem's solve is the
// same pass sequence as the single solve (5 passes for FP64; 5 per FP32
// solve / refinement for the mixed form), so step s of all problems is
// launched together: one launch per size class (problems sharing the cell
// count -- hence the launch configuration -- along the pass axis), one
// workgroup per (problem, tile) item of its work list, reading every
// problem's pass arguments from a device descriptor table.  The periodic
// copies follow as three launches (x, y, z faces, in that order, as in the
// single solve).  The whole sequence is built into one hipGraph at plan
// creation, the size classes of a step as parallel branches; a solve is one
// hipGraphLaunch.  The node arrays f of the problems
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
    int step;                 // pass index in the problems' sequences
};

// geometry of one problem of a plan (a rebind must present the same)
struct ProbKey {
    int N[3], bc[3];
    double dom[6];
    unsigned faces;   // bit i: face array i read by the codes
    bool singular;
    bool operator==(const ProbKey &o) const
    {
        for (int a = 0; a < 3; a++)
            if (N[a] != o.N[a] || bc[a] != o.bc[a]) return false;
        for (int i = 0; i < 6; i++)
            if (dom[i] != o.dom[i]) return false;
        return faces == o.faces && singular == o.singular;
    }
};

typedef std::tuple<int, int, int, int, unsigned long long> LaunchKey;   // N, ept, TPL, L, plan

// step s of the problems sharing one launch configuration: one launch
struct GroupRec {
    int step;
    LaunchKey key;
    std::vector<int> probs;
    size_t d_off;   // descriptors in the table image
};

struct BatchPlan {
    int dev = 0, count = 0, prec = 64, refine = 0;   // refine < 0: per-problem counts up to -refine
    std::vector<int> kref;                            // refinement steps of each problem
    void *work = nullptr;     // per-problem workspaces
    void *tables = nullptr;   // device image: descriptors, periodic copies | items | twiddles, eigenvalues
    size_t work_bytes = 0, table_bytes = 0;
    size_t desc_bytes = 0;                // descriptors and copy lists: the part a rebind rewrites
    std::vector<char> host;               // host image of the tables
    char *staging = nullptr;              // pinned copy of the rewritable part (rebind uploads)
    std::vector<GroupRec> groups;
    size_t c_off[3] = {0, 0, 0};          // periodic copy lists in the image
    std::vector<size_t> ws_off, u_off;    // workspace / FP64 iterate offsets in work
    std::vector<size_t> u32_off;          // FP32 first iterate offsets (mixed solves on lean-I/O kernels)
    std::vector<char> u32;                // problem keeps its first iterate in FP32
    // the plan's own twiddle and eigenvalue tables (image offsets): the
    // descriptors point here, never at the library's caches
    std::map<std::pair<int, int>, size_t> tw_off;   // (N, precision)
    std::vector<size_t> eig_off;                    // per problem
    std::vector<ProbKey> keys;
    std::vector<BatchLaunch> launches;
    const BatchCopy *copies[3] = {nullptr, nullptr, nullptr};
    int ncopies[3] = {0, 0, 0}, copy_grid[3] = {0, 0, 0};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // stream order across the streams a plan is used on: `done` marks the
    // end of the last solve or upload enqueued for the plan; `uploaded` the
    // last rebind upload (the staging buffer is free again once it has run)
    std::mutex mu;
    hipEvent_t done = nullptr, uploaded = nullptr;
    bool done_rec = false, up_rec = false;
    ~BatchPlan()
    {
        if (done && done_rec) (void)hipEventSynchronize(done);   // the plan may still run on some stream
        if (uploaded && up_rec) (void)hipEventSynchronize(uploaded);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (tables) (void)hipFree(tables);
        if (work) (void)hipFree(work);
        if (staging) (void)hipHostFree(staging);
        if (done) (void)hipEventDestroy(done);
        if (uploaded) (void)hipEventDestroy(uploaded);
    }
    // order st after the plan's previous use (no host wait)
    hipError_t wait_on(hipStream_t st) { return done_rec ? hipStreamWaitEvent(st, done, 0) : hipSuccess; }
    hipError_t mark(hipStream_t st)
    {
        hipError_t e = hipEventRecord(done, st);
        done_rec = done_rec || e == hipSuccess;
        return e;
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

// The launch configuration of a pass (problems of every kind share a launch:
// splitting the groups by kind, to run kernels that hold one kind's code,
// turned the 2048-problem workload's 28 launches into 128 and measured
// slower, FP64 20.8 -> 22.2 ms)
static LaunchKey launch_key(const PassSpec &q)
{
    return std::make_tuple(q.N, q.ept, q.d.p.TPL, q.d.p.L, q.d.p.plan);
}

static ProbKey key_of(const SolveGeom &g, const chz_problem &q)
{
    ProbKey k{};
    for (int a = 0; a < 3; a++) {
        k.N[a] = g.N[a];
        k.bc[a] = g.bc[a];
    }
    for (int i = 0; i < 6; i++) {
        k.dom[i] = q.domain[i];
        if (g.faces[i]) k.faces |= 1u << i;
    }
    k.singular = g.singular;
    return k;
}

// pass sequence of every problem (the single solve's builder) on the plan's workspaces
static int build_specs(Context &ctx, const chz_problem *probs, const BatchPlan &plan, const std::vector<SolveGeom> &geo,
                       std::vector<std::vector<PassSpec>> &specs)
{
    char *wbase = reinterpret_cast<char *>(plan.work);
    specs.assign(plan.count, {});
    for (int i = 0; i < plan.count; i++) {
        const int refine = plan.kref[i];
        const SolveGeom &g = geo[i];
        double *f = probs[i].f;
        std::vector<PassSpec> &v = specs[i];
        int rc;
        if (plan.prec == 64) {
            double *ex = reinterpret_cast<double *>(wbase + plan.ws_off[i]);
            if (!(rc = solve_passes<double>(ctx, g, f, ex, nullptr, nullptr, v))) rc = final_pass<double>(ctx, g, ex, nullptr, f, v);
        } else {
            // as helmholtz_mixed: the first iterate in FP32 where the plan keeps it so
            float *ex = reinterpret_cast<float *>(wbase + plan.ws_off[i]);
            const bool u32 = plan.u32[i] != 0;
            double *U = refine > 0 && (!u32 || refine > 1) ? reinterpret_cast<double *>(wbase + plan.u_off[i]) : nullptr;
            double *U0 = u32 ? reinterpret_cast<double *>(wbase + plan.u32_off[i]) : nullptr;   // FP32 data
            if (!(rc = solve_passes<float>(ctx, g, f, ex, nullptr, nullptr, v)))
                rc = final_pass<float>(ctx, g, ex, nullptr, refine == 0 ? f : u32 ? U0 : U, v, nullptr, u32 ? 2 : 0);
            for (int it = 1; it <= refine && rc == OK; it++) {
                const bool from32 = u32 && it == 1;
                const double *Uin = from32 ? U0 : U;
                if (!(rc = solve_passes<float>(ctx, g, f, ex, Uin, nullptr, v, nullptr, from32 ? 1 : 0)))
                    rc = final_pass<float>(ctx, g, ex, Uin, it == refine ? f : U, v, nullptr, from32 ? 1 : 0);
            }
        }
        if (rc) return fail(rc, "problem " + std::to_string(i) + ": " + last_error());
    }
    return OK;
}

// descriptors of every group and the periodic copy lists into the table
// image; the twiddle and eigenvalue pointers of the descriptors are redirected
// to the plan's own copies (device base of the plan's table allocation)
static void write_descs(BatchPlan &plan, const chz_problem *probs, const std::vector<SolveGeom> &geo,
                        const std::vector<std::vector<PassSpec>> &specs, char *image)
{
    char *tdev = reinterpret_cast<char *>(plan.tables);
    for (const GroupRec &gr : plan.groups) {
        BatchDesc *d = reinterpret_cast<BatchDesc *>(image + gr.d_off);
        for (size_t k = 0; k < gr.probs.size(); k++) {
            const int i = gr.probs[k];
            const PassSpec &q = specs[i][gr.step];
            d[k] = q.d;
            d[k].p.tw = tdev + plan.tw_off.at(std::make_pair(q.N, q.prec));
            if (d[k].sa.eig) d[k].sa.eig = reinterpret_cast<const double *>(tdev + plan.eig_off[i]);
        }
    }
    for (int a = 0; a < 3; a++) {
        BatchCopy *c = reinterpret_cast<BatchCopy *>(image + plan.c_off[a]);
        int n = 0;
        for (int i = 0; i < plan.count; i++)
            if (geo[i].bc[a] == 0) c[n++] = BatchCopy{probs[i].f, geo[i].N[0], geo[i].N[1], geo[i].N[2], a};
    }
}

// Refinement steps of one problem when the plan chooses them (refine < 0).
// The error of the FP32 solve after k steps grows with the spectral condition
// number kappa = max|e + lambda| / min|e + lambda|: on the 2048-problem
// workload (sizes 32..128, codes 0-4, lambda in [-10, 0.5]) one step left
// every problem with kappa <= 1e5 within 4e-11 of the FP64 solve, and up to
// 1e-9 above that.  One step up to kappa = 1e5, two up to 1e7, three beyond;
// at most kmax.
static int auto_refine(const SolveGeom &g, int kmax)
{
    const double kappa = spectral_condition(g);
    const int k = kappa <= 1e5 ? 1 : kappa <= 1e7 ? 2 : 3;
    return std::min(k, kmax);
}

static int create(const chz_problem *probs, int count, int prec, int refine, BatchPlan **out)
{
    if (!out) return fail(ERR_INVALID, "null plan pointer");
    *out = nullptr;
    if (!probs || count <= 0) return fail(ERR_INVALID, "empty problem list");
    if (prec != 64 && prec != 32) return fail(ERR_INVALID, "precision must be 64 or 32");
    if (prec == 64 && refine != 0) return fail(ERR_INVALID, "the FP64 batched solve takes no refinement steps");
    if (refine < -16 || refine > 16) return fail(ERR_INVALID, "refinement count must be in -16..16");
    Context &ctx = Context::get();
    std::lock_guard<std::recursive_mutex> lock(ctx.mutex());

    std::vector<SolveGeom> geo(count);
    const size_t esz = prec == 64 ? 8 : 4;
    std::vector<size_t> ws_off(count), u_off(count, 0), u32_off(count, 0);
    std::vector<char> u32(count, 0);
    std::vector<int> kref(count, refine);
    size_t total = 0;
    for (int i = 0; i < count; i++) {
        const chz_problem &q = probs[i];
        if (!q.f) return fail(ERR_INVALID, "problem " + std::to_string(i) + ": null solution array");
        int rc = make_geom(geo[i], q.domain, q.N, q.bc, q.lambda, q.faces);
        if (rc) return fail(rc, "problem " + std::to_string(i) + ": " + last_error());
        if (refine < 0) kref[i] = auto_refine(geo[i], -refine);
        if (prec == 32 && kref[i] > 0 && geo[i].singular)
            return fail(ERR_UNSUPPORTED, "problem " + std::to_string(i) +
                                             ": singular problems (lambda = 0, periodic/Neumann-Neumann on every axis) "
                                             "are solved in FP64 or alone in mixed precision");
        ws_off[i] = total;
        total += align256(esz * (size_t)ws_row(q.N[0], esz) * (q.N[1] + 2) * (q.N[2] + 2));
        const size_t nodes = (size_t)(q.N[0] + 1) * (q.N[1] + 1) * (q.N[2] + 1);
        if (prec == 32 && kref[i] > 0 && iterate_f32_ok(ctx, geo[i], true)) {
            u32[i] = 1;
            u32_off[i] = total;
            total += align256(sizeof(float) * nodes);
        }
        if (prec == 32 && kref[i] > 0 && (!u32[i] || kref[i] > 1)) {
            u_off[i] = total;
            total += align256(sizeof(double) * nodes);
        }
    }
    BatchPlan *plan = new BatchPlan();
    plan->count = count;
    plan->prec = prec;
    plan->refine = refine;
    plan->kref = kref;
    plan->dev = ctx.device();
    plan->work_bytes = total;
    plan->ws_off = ws_off;
    plan->u_off = u_off;
    plan->u32_off = u32_off;
    plan->u32 = u32;
    for (int i = 0; i < count; i++) plan->keys.push_back(key_of(geo[i], probs[i]));
    if (hipMalloc(&plan->work, total) != hipSuccess) {
        delete plan;
        return fail(ERR_ALLOC, "batched workspace allocation failed (" + std::to_string(total) + " bytes)");
    }

    // pass sequence of every problem
    std::vector<std::vector<PassSpec>> specs;
    int rc = build_specs(ctx, probs, *plan, geo, specs);
    if (rc) {
        delete plan;
        return rc;
    }
    // problems with fewer refinement steps have shorter sequences: step s
    // groups the problems that have one
    size_t nsteps = 0;
    for (int i = 0; i < count; i++) nsteps = std::max(nsteps, specs[i].size());

    // group step s of all problems by launch configuration
    std::vector<GroupRec> &groups = plan->groups;
    for (size_t s = 0; s < nsteps; s++) {
        std::map<LaunchKey, size_t> idx;
        for (int i = 0; i < count; i++) {
            if (s >= specs[i].size()) continue;
            const PassSpec &q = specs[i][s];
            const LaunchKey key = launch_key(q);
            auto it = idx.find(key);
            if (it == idx.end()) {
                idx[key] = groups.size();
                groups.push_back(GroupRec{(int)s, key, {i}, 0});
            } else {
                groups[it->second].probs.push_back(i);
            }
        }
    }
    // table layout: descriptors of every group and the periodic copy lists
    // (rewritten by a rebind), then the work items, then the plan's own
    // twiddle and eigenvalue tables
    std::vector<size_t> i_off(groups.size());
    std::vector<int> nitems(groups.size());
    size_t tb = 0;
    for (size_t gi = 0; gi < groups.size(); gi++) {
        GroupRec &gr = groups[gi];
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
        gr.d_off = tb;
        tb += align256(sizeof(BatchDesc) * gr.probs.size());
    }
    for (int a = 0; a < 3; a++) {
        int maxface = 0, n = 0;
        for (int i = 0; i < count; i++) {
            const SolveGeom &g = geo[i];
            if (g.bc[a] != 0) continue;
            n++;
            const int face = a == 0 ? (g.N[1] + 1) * (g.N[2] + 1) : a == 1 ? (g.N[0] + 1) * (g.N[2] + 1)
                                                                             : (g.N[0] + 1) * (g.N[1] + 1);
            maxface = std::max(maxface, face);
        }
        if (n > 65535) {
            delete plan;
            return fail(ERR_INVALID, "more than 65535 periodic problems along one axis");
        }
        plan->ncopies[a] = n;
        plan->copy_grid[a] = std::min(64, (maxface + 255) / 256);
        plan->c_off[a] = tb;
        tb += align256(sizeof(BatchCopy) * (size_t)n);
    }
    plan->desc_bytes = tb;
    for (size_t gi = 0; gi < groups.size(); gi++) {
        i_off[gi] = tb;
        tb += align256(sizeof(BatchItem) * (size_t)nitems[gi]);
    }
    // twiddles per (N, precision) and eigenvalues per distinct geometry
    std::vector<std::pair<size_t, std::vector<char>>> blobs;
    for (int i = 0; i < count; i++)
        for (const PassSpec &q : specs[i]) {
            const auto key = std::make_pair(q.N, q.prec);
            if (plan->tw_off.count(key)) continue;
            plan->tw_off[key] = tb;
            blobs.push_back({tb, twiddle_bytes(q.N, q.prec)});
            tb += align256(blobs.back().second.size());
        }
    {
        std::map<std::string, size_t> seen;
        plan->eig_off.assign(count, 0);
        for (int i = 0; i < count; i++) {
            const std::vector<double> t = eig_tables(geo[i]);
            std::string k(reinterpret_cast<const char *>(t.data()), t.size() * sizeof(double));
            auto it = seen.find(k);
            if (it != seen.end()) {
                plan->eig_off[i] = it->second;
                continue;
            }
            plan->eig_off[i] = seen[k] = tb;
            blobs.push_back({tb, std::vector<char>(k.begin(), k.end())});
            tb += align256(k.size());
        }
    }
    plan->table_bytes = tb;
    if (hipMalloc(&plan->tables, tb) != hipSuccess) {
        delete plan;
        return fail(ERR_ALLOC, "batched descriptor table allocation failed");
    }
    if (hipHostMalloc((void **)&plan->staging, plan->desc_bytes, hipHostMallocDefault) != hipSuccess) {
        plan->staging = nullptr;
        delete plan;
        return fail(ERR_ALLOC, "batched descriptor staging allocation failed");
    }
    if (hipEventCreateWithFlags(&plan->done, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&plan->uploaded, hipEventDisableTiming) != hipSuccess) {
        delete plan;
        return fail(ERR_DEVICE, "batched plan event creation failed");
    }
    std::vector<char> &host = plan->host;
    host.assign(tb, 0);
    for (const auto &b : blobs) memcpy(host.data() + b.first, b.second.data(), b.second.size());
    write_descs(*plan, probs, geo, specs, host.data());
    for (size_t gi = 0; gi < groups.size(); gi++) {
        const GroupRec &gr = groups[gi];
        BatchItem *it = reinterpret_cast<BatchItem *>(host.data() + i_off[gi]);
        int w = 0;
        for (size_t k = 0; k < gr.probs.size(); k++) {
            const PassSpec &q = specs[gr.probs[k]][gr.step];
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
    if (hipMemcpy(plan->tables, host.data(), tb, hipMemcpyHostToDevice) != hipSuccess) {
        delete plan;
        return fail(ERR_DEVICE, "batched descriptor table upload failed");
    }
    std::vector<char>().swap(host);   // a rebind writes the descriptors into the staging buffer
    char *tbase = reinterpret_cast<char *>(plan->tables);
    for (int a = 0; a < 3; a++) plan->copies[a] = reinterpret_cast<const BatchCopy *>(tbase + plan->c_off[a]);

    for (size_t gi = 0; gi < groups.size(); gi++) {
        const GroupRec &gr = groups[gi];
        const PassSpec &q = specs[gr.probs[0]][gr.step];
        BatchLaunch l;
        l.lds = axis_pass_lds(q.prec, q.d.p.M, q.d.p.L, q.d.p.TPL);
        // lean tile I/O where the windows of every problem of the group allow it
        // a kind-specialised kernel where every problem of the group has the same kind
        int ks = pass_kind_sel(q.op, q.d.p, &q.d.sa);
        for (size_t k = 1; ks >= 0 && k < gr.probs.size(); k++) {
            const PassSpec &qk = specs[gr.probs[k]][gr.step];
            if (pass_kind_sel(qk.op, qk.d.p, &qk.d.sa) != ks) ks = -1;
        }
        const FixedEntry *ln = batch_line_kernel(q.prec, q.op, q.N, q.ept, q.d.p.TPL, q.d.p.L, q.d.p.plan, ks);
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
        for (size_t k = 0; !ln && k < gr.probs.size(); k++)
            if (specs[gr.probs[k]][gr.step].d.p.u_f32) {
                delete plan;
                return fail(ERR_INVALID, "internal: FP32 iterate on a pass without a lean-I/O kernel");
            }
        l.threads = q.d.p.L * q.d.p.TPL;
        // one workgroup per (problem, tile) item, in whole XCD rounds (item chunks
        // per XCD): workgroups that start and end independently overlap one
        // tile's memory phase with another's compute better than a persistent
        // grid of resident workgroups walking the items (2048-problem workload:
        // FP64 22.6 -> 20.9 ms, FP32 + refinement 40.1 -> 34.2 ms)
        long long grid = (nitems[gi] + 7) / 8 * 8;
        l.grid = (int)std::max<long long>(grid, 1);
        l.items = reinterpret_cast<const BatchItem *>(tbase + i_off[gi]);
        l.nitems = nitems[gi];
        l.descs = reinterpret_cast<const BatchDesc *>(tbase + gr.d_off);
        l.name = std::string(q.name) + "/" + std::to_string(q.N);
        l.step = gr.step;
        if (l.nitems > 0) plan->launches.push_back(l);
    }

    // the graph: the launches of one step (one per size class, each over its
    // own problems) are independent nodes that wait for every launch of the
    // step before; the periodic copies follow, x, y, z faces in turn
    hipGraph_t graph = nullptr;
    hipError_t e = hipGraphCreate(&graph, 0);
    plan->graph = graph;
    std::vector<hipGraphNode_t> prev, cur;
    int cur_step = -1;
    for (const BatchLaunch &l : plan->launches) {
        if (e != hipSuccess) break;
        if (l.step != cur_step) {
            if (!cur.empty()) prev.swap(cur);
            cur.clear();
            cur_step = l.step;
        }
        const BatchItem *items = l.items;
        int n = l.nitems;
        const BatchDesc *descs = l.descs;
        void *args[] = {&items, &n, &descs};
        hipKernelNodeParams kp{};
        kp.func = const_cast<void *>(l.fn);
        kp.gridDim = dim3((unsigned)l.grid);
        kp.blockDim = dim3((unsigned)l.threads);
        kp.sharedMemBytes = (unsigned)l.lds;
        kp.kernelParams = args;
        kp.extra = nullptr;
        hipGraphNode_t node = nullptr;
        e = hipGraphAddKernelNode(&node, graph, prev.data(), prev.size(), &kp);
        if (e == hipSuccess) cur.push_back(node);
    }
    for (int a = 0; a < 3 && e == hipSuccess; a++) {
        if (!plan->ncopies[a]) continue;
        const BatchCopy *tab = plan->copies[a];
        void *args[] = {&tab};
        hipKernelNodeParams kp{};
        kp.func = const_cast<void *>(batch_periodic_kernel());
        kp.gridDim = dim3((unsigned)plan->copy_grid[a], (unsigned)plan->ncopies[a]);
        kp.blockDim = dim3(256);
        kp.sharedMemBytes = 0;
        kp.kernelParams = args;
        kp.extra = nullptr;
        hipGraphNode_t node = nullptr;
        e = hipGraphAddKernelNode(&node, graph, cur.data(), cur.size(), &kp);
        if (e == hipSuccess) cur.assign(1, node);
    }
    if (e != hipSuccess) {
        delete plan;
        return fail(ERR_DEVICE, std::string("graph build: ") + hipGetErrorString(e));
    }
    if ((e = hipGraphInstantiate(&plan->exec, graph, nullptr, nullptr, 0)) != hipSuccess) {
        delete plan;
        return fail(ERR_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(e));
    }
    *out = plan;
    return OK;
}

// Point a plan at a new problem list of the same geometry (sizes, codes,
// domains, which face arrays, singularity): new f / face arrays and lambdas.
// The descriptors and periodic copy lists are rebuilt on the host and
// uploaded on st from the plan's pinned staging buffer, ordered after the
// plan's earlier solves on any stream (the items, tables, workspaces and graph
// stay).  The host waits only for the plan's previous upload to have read the
// staging buffer.  *cfg_changed is set when the launch configurations changed
// since the plan was made (tuning calls): the plan must then be rebuilt.
static int rebind(BatchPlan &plan, const chz_problem *probs, int count, hipStream_t st, bool *cfg_changed = nullptr)
{
    if (cfg_changed) *cfg_changed = false;
    if (!probs || count != plan.count) return fail(ERR_INVALID, "rebind: the problem count differs from the plan's");
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev != plan.dev) return fail(ERR_INVALID, "plan belongs to another device");
    Context &ctx = Context::get();
    std::lock_guard<std::recursive_mutex> lock(ctx.mutex());
    std::vector<SolveGeom> geo(count);
    for (int i = 0; i < count; i++) {
        const chz_problem &q = probs[i];
        if (!q.f) return fail(ERR_INVALID, "problem " + std::to_string(i) + ": null solution array");
        int rc = make_geom(geo[i], q.domain, q.N, q.bc, q.lambda, q.faces);
        if (rc) return fail(rc, "problem " + std::to_string(i) + ": " + last_error());
        if (!(key_of(geo[i], q) == plan.keys[i]))
            return fail(ERR_INVALID, "rebind: problem " + std::to_string(i) + " differs in geometry from the plan's");
        if (plan.refine < 0 && auto_refine(geo[i], -plan.refine) != plan.kref[i]) {
            if (cfg_changed) *cfg_changed = true;
            return fail(ERR_INVALID, "rebind: the new lambda of problem " + std::to_string(i) +
                                         " changes its refinement count; make a new plan");
        }
    }
    std::vector<std::vector<PassSpec>> specs;
    int rc = build_specs(ctx, probs, plan, geo, specs);
    if (rc) return rc;
    for (const GroupRec &gr : plan.groups)
        for (int i : gr.probs) {
            const PassSpec &q = specs[i][gr.step];
            if (launch_key(q) != gr.key) {
                if (cfg_changed) *cfg_changed = true;
                return fail(ERR_INVALID, "rebind: launch configurations changed since the plan was made");
            }
        }
    std::lock_guard<std::mutex> plock(plan.mu);
    hipError_t e = hipSuccess;
    if (plan.up_rec) e = hipEventSynchronize(plan.uploaded);   // staging buffer free again
    if (e == hipSuccess) {
        write_descs(plan, probs, geo, specs, plan.staging);
        e = plan.wait_on(st);   // earlier solves of the plan have read the old descriptors
    }
    if (e == hipSuccess) e = hipMemcpyAsync(plan.tables, plan.staging, plan.desc_bytes, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        e = hipEventRecord(plan.uploaded, st);
        plan.up_rec = plan.up_rec || e == hipSuccess;
    }
    if (e == hipSuccess) e = plan.mark(st);
    if (e != hipSuccess) return fail(ERR_DEVICE, std::string("rebind upload: ") + hipGetErrorString(e));
    return OK;
}

// Replay the plan's graph (or issue its launches) on st after its previous use.
static int launch_plan(Context &ctx, BatchPlan &plan, bool use_graph, hipStream_t st)
{
    std::lock_guard<std::mutex> plock(plan.mu);
    hipError_t e = plan.wait_on(st);
    if (e != hipSuccess) return fail(ERR_DEVICE, std::string("plan stream wait: ") + hipGetErrorString(e));
    int rc = OK;
    if (use_graph) {
        e = hipGraphLaunch(plan.exec, st);
        if (e != hipSuccess) rc = fail(ERR_DEVICE, std::string("graph launch: ") + hipGetErrorString(e));
    } else {
        std::lock_guard<std::recursive_mutex> lock(ctx.mutex());
        const bool tm = ctx.timing();
        rc = enqueue(ctx, plan, st, tm);
        if (tm) ctx.timing_end(st);
    }
    if ((e = plan.mark(st)) != hipSuccess && rc == OK) rc = fail(ERR_DEVICE, std::string("plan event: ") + hipGetErrorString(e));
    return rc;
}

// Plans kept for repeated solves of one geometry (chz_batch_solve_cached):
// up to kCached per process, least recently used first out.  Never freed at
// exit (the device may be gone by then).
static const size_t kCached = 4;
struct PlanCache {
    std::mutex mu;
    std::vector<std::pair<std::string, BatchPlan *>> plans;   // most recent last
};
static PlanCache &plan_cache()
{
    static PlanCache *c = new PlanCache();
    return *c;
}

// drop every cached plan (chz_release); each waits for its own last use
void release_cached_plans()
{
    PlanCache &c = plan_cache();
    std::lock_guard<std::mutex> lock(c.mu);
    for (auto &e : c.plans) delete e.second;
    c.plans.clear();
}

static std::string signature(const chz_problem *probs, int count, int prec, int refine, int dev)
{
    std::string s;
    auto put = [&s](const void *p, size_t n) { s.append(reinterpret_cast<const char *>(p), n); };
    put(&dev, sizeof dev);
    put(&prec, sizeof prec);
    put(&refine, sizeof refine);
    put(&count, sizeof count);
    for (int i = 0; i < count; i++) {
        const chz_problem &q = probs[i];
        put(q.N, sizeof q.N);
        put(q.bc, sizeof q.bc);
        put(q.domain, sizeof q.domain);
        unsigned faces = 0;
        for (int k = 0; k < 6; k++)
            if (q.faces[k]) faces |= 1u << k;
        put(&faces, sizeof faces);
        const char zero = q.lambda == 0.0;   // singularity also depends on lambda = 0
        put(&zero, 1);
    }
    return s;
}

static int solve_cached(const chz_problem *probs, int count, int prec, int refine, hipStream_t st)
{
    if (!probs || count <= 0) return fail(ERR_INVALID, "empty problem list");
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return fail(ERR_DEVICE, "no current device");
    const std::string sig = signature(probs, count, prec, refine, dev);
    // lock order: the context, then the cache (callers holding the context re-enter it)
    Context &ctx = Context::get();
    std::lock_guard<std::recursive_mutex> clock(ctx.mutex());
    PlanCache &c = plan_cache();
    std::lock_guard<std::mutex> lock(c.mu);
    BatchPlan *plan = nullptr;
    for (size_t i = 0; i < c.plans.size(); i++)
        if (c.plans[i].first == sig) {
            plan = c.plans[i].second;
            c.plans.erase(c.plans.begin() + (long)i);
            break;
        }
    int rc;
    if (plan) {
        bool cfg_changed = false;
        if ((rc = rebind(*plan, probs, count, st, &cfg_changed))) {
            delete plan;   // waits for its last use
            plan = nullptr;
            if (!cfg_changed) return rc;
        }
    }
    if (!plan) {
        if ((rc = create(probs, count, prec, refine, &plan))) return rc;
        if (c.plans.size() >= kCached) {
            delete c.plans.front().second;   // waits for its last use
            c.plans.erase(c.plans.begin());
        }
    }
    c.plans.push_back({sig, plan});
    return launch_plan(ctx, *plan, true, st);
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
    return chz::launch_plan(chz::Context::get(), *p, use_graph != 0, (hipStream_t)stream);
}

int chz_batch_rebind(void *plan, const chz_problem *problems, int count, void *stream)
{
    if (!plan) return chz::fail(chz::ERR_INVALID, "null plan");
    return chz::rebind(*reinterpret_cast<BatchPlan *>(plan), problems, count, (hipStream_t)stream);
}

int chz_batch_solve_cached(const chz_problem *problems, int count, int precision, int refine, void *stream)
{
    return chz::solve_cached(problems, count, precision, refine, (hipStream_t)stream);
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

int chz_batch_refine_counts(void *plan, int *counts, int count)
{
    if (!plan || !counts) return chz::fail(chz::ERR_INVALID, "null plan or counts");
    const BatchPlan *p = reinterpret_cast<const BatchPlan *>(plan);
    if (count != p->count) return chz::fail(chz::ERR_INVALID, "count differs from the plan's problem count");
    for (int i = 0; i < count; i++) counts[i] = p->kref[(size_t)i];
    return chz::OK;
}

int chz_batch_destroy(void *plan)
{
    delete reinterpret_cast<BatchPlan *>(plan);
    return chz::OK;
}

}  // extern "C"
