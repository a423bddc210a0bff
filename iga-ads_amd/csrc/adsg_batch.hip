// Batches of independent 2D problems advanced together (declared in include/adsgpu.h).
//
// Every problem keeps its own 1D tables and FP64-computed factors (an adsg_sim without state
// vectors), while the state of all problems lives in shared arenas, problem after problem.
// One step is a fixed sequence of grouped launches over all problems: zero fill of the
// arena, per degree p and parity colour one element-kernel launch whose blocks cover the
// tiles of every problem of that degree, per degree one x-sweep and one y-sweep launch over
// all their lines, and one elementwise update of the whole arena. The step sequences are
// replayed from captured HIP graphs of at most 50 steps. The problems' host set-up (tables,
// band LU, segment coefficients) runs on worker threads and stages its uploads, which then
// go to the device as one arena per phase (create, prepare) with one copy each.
#include <hip/hip_runtime.h>

#include <atomic>
#include <thread>

#include <algorithm>
#include <cstring>
#include <map>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "adsg_sim.hpp"

namespace {

thread_local std::string g_batch_error;

// the launches of one degree for one arena parity (device arrays of items)
struct batch_group {
    int p = 0;
    int nitems = 0;
    void* rhs_items = nullptr;  // rhs2_item<T>[nitems]
    int rhs_blocks[4] = {0, 0, 0, 0};
    void* sx_items = nullptr;  // sweep_item<T>[nitems], x lines
    int sx_blocks = 0;
    void* sy_items = nullptr;  // y lines
    int sy_blocks = 0;
    // segmented sweeps (when every problem's lines split into segments)
    bool seg = false;
    void* gx_items = nullptr;  // seg_item<T>[nitems], x lines
    int gx_blocks = 0;
    void* gy_items = nullptr;
    int gy_blocks = 0;
};

}  // namespace

struct adsg_batch {
    int precision = ADSG_FP64;
    hipStream_t stream = nullptr;
    std::vector<adsg_sim*> sims;
    std::vector<long> off;  // first coefficient of each problem in the arenas
    long total = 0;
    // state arenas: u[phase] holds u, u[1 - phase] holds u_prev (FP64 in both precisions);
    // FP32: sh[.] their FP32 copies and buf the increment
    double* u[2] = {nullptr, nullptr};
    float* sh[2] = {nullptr, nullptr};
    float* buf = nullptr;
    int phase = 0;
    std::vector<adsg_rhs_form> forms;
    bool prepared = false;
    int shape = 0;  // element tile shape: 0 = 16 x 16, 1 = 8 x 8
    int sweep = 0;  // 0: segmented line sweeps where every problem allows, 1: one lane per line
    std::vector<batch_group> groups[2];
    void* packs = nullptr;  // packed sweep rows of every problem and direction (step precision)
    std::map<std::pair<int, int>, hipGraphExec_t> graphs;  // (nsteps, phase at start)
    // every problem's tables (create) and factors (prepare), each set in one device arena
    void* arena[2] = {nullptr, nullptr};
    std::string err;
};

namespace {

// f(0) .. f(count - 1) on up to 16 host threads (the problems' set-up is independent)
template <typename F>
void for_each_problem(int count, F&& f) {
    const int hw = static_cast<int>(std::max(1u, std::thread::hardware_concurrency()));
    const int nt = std::max(1, std::min({16, hw, count}));
    int dev = 0;
    (void)hipGetDevice(&dev);  // the workers use the caller's device (HIP's is per thread)
    std::atomic<int> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t)
        pool.emplace_back([&] {
            (void)hipSetDevice(dev);
            for (int i = next++; i < count; i = next++) f(i);
        });
    for (auto& th : pool) th.join();
}

int bfail(adsg_batch* b, int code, const std::string& msg) {
    if (b) b->err = msg;
    return code;
}

int bhip(adsg_batch* b, hipError_t e, const char* what) {
    return bfail(b, ADSG_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

void free_groups(adsg_batch* b) {
    for (auto& gs : b->groups) {
        for (auto& g : gs)
            for (void* p : {g.rhs_items, g.sx_items, g.sy_items, g.gx_items, g.gy_items})
                if (p) (void)hipFree(p);
        gs.clear();
    }
    if (b->packs) (void)hipFree(b->packs);
    b->packs = nullptr;
}

void drop_batch_graphs(adsg_batch* b) {
    for (auto& kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
    b->graphs.clear();
}

template <typename T>
adsg::dev_tables<T> tables_of(const adsg_sim* h, int d) {
    adsg::dev_tables<T> t;
    t.val = static_cast<const T*>(h->dirs[d].val);
    t.der = static_cast<const T*>(h->dirs[d].der);
    t.wj = static_cast<const T*>(h->dirs[d].wj);
    t.elements = h->dirs[d].tab.elements;
    t.dofs = h->dirs[d].tab.dofs;
    return t;
}

template <typename T>
adsg::dev_factors<T> factors_of(const adsg_sim* h, int d) {
    const adsg::dev_slot& s = h->dirs[d].slots[0];
    adsg::dev_factors<T> f;
    f.lower = static_cast<const T*>(s.lower);
    f.upper = static_cast<const T*>(s.upper);
    f.rdiag = static_cast<const T*>(s.rdiag);
    f.poff = s.poff;
    f.n = s.lf.n;
    f.kl = s.lf.kl;
    f.ku = s.lf.ku_eff;
    f.pivoted = s.lf.pivoted;
    return f;
}

template <typename T>
int upload_items(adsg_batch* b, void** dst, const std::vector<T>& items) {
    if (items.empty()) return ADSG_OK;
    if (hipMalloc(dst, sizeof(T) * items.size()) != hipSuccess)
        return bfail(b, ADSG_ENOMEM, "batch: no memory for the launch items");
    hipError_t e = hipMemcpy(*dst, items.data(), sizeof(T) * items.size(), hipMemcpyHostToDevice);
    return e == hipSuccess ? ADSG_OK : bhip(b, e, "batch: uploading launch items");
}

// launch items of every degree for both arena parities (T: the step precision)
template <typename T>
int build_groups(adsg_batch* b) {
    free_groups(b);
    drop_batch_graphs(b);
    // batched segmentation of every problem's two Gram matrices, packed rows in one arena
    const size_t P = b->sims.size();
    std::vector<adsg::line_segments> segs(2 * P);
    std::vector<size_t> pack_off(2 * P, 0);
    bool all_seg = true;
    size_t total_pack = 0;
    for (size_t i = 0; i < P && all_seg; ++i)
        for (int d = 0; d < 2; ++d) {
            auto& sg = segs[2 * i + d];
            if (!adsg::segment_lines(b->sims[i]->dirs[d].slots[0].lf, sg, 0, adsg::kBatchSegments)) {
                all_seg = false;
                break;
            }
            pack_off[2 * i + d] = total_pack;
            total_pack += sg.pack.size();
        }
    if (all_seg) {
        std::vector<T> arena(std::max<size_t>(total_pack, 1));
        for (size_t k = 0; k < 2 * P; ++k)
            for (size_t j = 0; j < segs[k].pack.size(); ++j) arena[pack_off[k] + j] = static_cast<T>(segs[k].pack[j]);
        if (hipMalloc(&b->packs, sizeof(T) * arena.size()) != hipSuccess)
            return bfail(b, ADSG_ENOMEM, "batch: no memory for the sweep rows");
        hipError_t e = hipMemcpy(b->packs, arena.data(), sizeof(T) * arena.size(), hipMemcpyHostToDevice);
        if (e != hipSuccess) return bhip(b, e, "batch: uploading sweep rows");
    }
    for (int ph = 0; ph < 2; ++ph) {
        for (int p = 1; p <= 3; ++p) {
            std::vector<adsg::rhs2_item<T>> ri;
            std::vector<adsg::sweep_item<T>> sx, sy;
            std::vector<adsg::seg_item<T>> gx, gy;
            batch_group g;
            g.p = p;
            for (size_t i = 0; i < b->sims.size(); ++i) {
                const adsg_sim* h = b->sims[i];
                if (h->dirs[0].tab.p != p) continue;
                const adsg_rhs_form& fm = b->forms[i];
                adsg::rhs2_item<T> it;
                // a step starting at parity ph reads u_prev = u[ph] and writes u[1 - ph]
                T* target;
                if constexpr (sizeof(T) == sizeof(float)) {
                    it.up = b->sh[ph] + b->off[i];
                    target = b->buf + b->off[i];
                } else {
                    it.up = b->u[ph] + b->off[i];
                    target = b->u[1 - ph] + b->off[i];
                }
                it.rhs = target;
                it.X = tables_of<T>(h, 0);
                it.Y = tables_of<T>(h, 1);
                // FP32 assembles the increment rhs - M u_prev: the mass term less one
                it.prm.mass = static_cast<T>(sizeof(T) == sizeof(float) ? fm.mass - 1.0 : fm.mass);
                it.prm.dt = static_cast<T>(fm.dt);
                it.prm.gx = static_cast<T>(fm.grad[0]);
                it.prm.gy = static_cast<T>(fm.grad[1]);
                it.prm.gz = static_cast<T>(fm.grad[2]);
                it.prm.reaction = static_cast<T>(fm.reaction);
                int count[4];
                adsg::rhs2_batch_tiles(b->shape, it.X.elements, it.Y.elements, it.ntx2, count);
                for (int c = 0; c < 4; ++c) {
                    it.first[c] = g.rhs_blocks[c];
                    g.rhs_blocks[c] += count[c];
                }
                ri.push_back(it);
                const long nx = h->dirs[0].tab.dofs, ny = h->dirs[1].tab.dofs;
                adsg::sweep_item<T> xi, yi;
                xi.u = yi.u = target;
                xi.g.n_lines = ny;
                xi.g.S = 1;
                xi.f = factors_of<T>(h, 0);
                xi.first = g.sx_blocks;
                g.sx_blocks += static_cast<int>((ny + 63) / 64);
                yi.g.n_lines = nx;
                yi.g.A = nx;
                yi.g.B = nx * ny;
                yi.g.S = nx;
                yi.f = factors_of<T>(h, 1);
                yi.first = g.sy_blocks;
                g.sy_blocks += static_cast<int>((nx + 63) / 64);
                sx.push_back(xi);
                sy.push_back(yi);
                if (all_seg) {
                    adsg::seg_item<T> a, c;
                    const auto& sgx = segs[2 * i];
                    const auto& sgy = segs[2 * i + 1];
                    a.u = c.u = target;
                    a.g.n_lines = ny;  // x line l starts at l * nx
                    a.g.A = 1;
                    a.g.B = nx;
                    a.g.S = 1;
                    a.pack = static_cast<const T*>(b->packs) + pack_off[2 * i];
                    a.n = static_cast<int>(nx);
                    a.m = sgx.m;
                    a.nseg = sgx.nseg;
                    a.first = g.gx_blocks;
                    g.gx_blocks += static_cast<int>((ny + 63) / 64);
                    c.g = yi.g;
                    c.pack = static_cast<const T*>(b->packs) + pack_off[2 * i + 1];
                    c.n = static_cast<int>(ny);
                    c.m = sgy.m;
                    c.nseg = sgy.nseg;
                    c.first = g.gy_blocks;
                    g.gy_blocks += static_cast<int>((nx + 63) / 64);
                    gx.push_back(a);
                    gy.push_back(c);
                }
            }
            if (ri.empty()) continue;
            g.nitems = static_cast<int>(ri.size());
            b->groups[ph].push_back(g);
            batch_group& G = b->groups[ph].back();
            int rc;
            if ((rc = upload_items(b, &G.rhs_items, ri))) return rc;
            if ((rc = upload_items(b, &G.sx_items, sx))) return rc;
            if ((rc = upload_items(b, &G.sy_items, sy))) return rc;
            G.seg = all_seg;
            if (all_seg) {
                if ((rc = upload_items(b, &G.gx_items, gx))) return rc;
                if ((rc = upload_items(b, &G.gy_items, gy))) return rc;
            }
        }
    }
    return ADSG_OK;
}

int build(adsg_batch* b) {
    return b->precision == ADSG_FP32 ? build_groups<float>(b) : build_groups<double>(b);
}

// one step of every problem, starting at arena parity b->phase
template <typename T>
int batch_step_impl(adsg_batch* b) {
    const int ph = b->phase;
    hipError_t e;
    T* zero = nullptr;
    if constexpr (sizeof(T) == sizeof(float))
        zero = b->buf;
    else
        zero = b->u[1 - ph];
    if ((e = adsg::launch_fill<T>(zero, b->total, T(0), b->stream)) != hipSuccess) return bhip(b, e, "batch: zero fill");
    for (const batch_group& g : b->groups[ph])
        for (int c = 0; c < 4; ++c) {
            e = adsg::launch_rhs2_batch<T>(g.p, b->shape, static_cast<const adsg::rhs2_item<T>*>(g.rhs_items),
                                          g.nitems, c, g.rhs_blocks[c], b->stream);
            if (e != hipSuccess) return bhip(b, e, "batch: element kernels");
        }
    for (const batch_group& g : b->groups[ph]) {
        if (g.seg && b->sweep == 0)
            e = adsg::launch_sweep_seg_batch<T>(g.p, g.p, static_cast<const adsg::seg_item<T>*>(g.gx_items),
                                                g.nitems, g.gx_blocks, b->stream);
        else
            e = adsg::launch_sweep_batch<T>(g.p, g.p, static_cast<const adsg::sweep_item<T>*>(g.sx_items),
                                            g.nitems, g.sx_blocks, b->stream);
        if (e != hipSuccess) return bhip(b, e, "batch: x sweeps");
    }
    for (const batch_group& g : b->groups[ph]) {
        if (g.seg && b->sweep == 0)
            e = adsg::launch_sweep_seg_batch<T>(g.p, g.p, static_cast<const adsg::seg_item<T>*>(g.gy_items),
                                                g.nitems, g.gy_blocks, b->stream);
        else
            e = adsg::launch_sweep_batch<T>(g.p, g.p, static_cast<const adsg::sweep_item<T>*>(g.sy_items),
                                            g.nitems, g.sy_blocks, b->stream);
        if (e != hipSuccess) return bhip(b, e, "batch: y sweeps");
    }
    if constexpr (sizeof(T) == sizeof(float)) {
        // u = u_prev + increment in FP64, with its FP32 copy (no boundary rows: natural BC)
        e = adsg::launch_apply_increment(b->u[ph], b->buf, b->u[1 - ph], b->sh[1 - ph], 2,
                                         static_cast<int>(b->total), 1, 1, false, b->stream);
        if (e != hipSuccess) return bhip(b, e, "batch: update");
    }
    b->phase = 1 - ph;
    return ADSG_OK;
}

int batch_step(adsg_batch* b) {
    return b->precision == ADSG_FP32 ? batch_step_impl<float>(b) : batch_step_impl<double>(b);
}

// Steps per captured graph: longer runs replay graphs of this many steps one after another
// (even, so the state arenas' parity is the same at every chunk start). It bounds the node
// count of one graph (about 20 kernels per step) and its instantiation cost.
constexpr int kMaxGraphSteps = 50;

int run_batch_chunk(adsg_batch* b, int nsteps);

int run_batch(adsg_batch* b, int nsteps) {
    while (nsteps > 0) {
        const int n = nsteps > kMaxGraphSteps ? kMaxGraphSteps : nsteps;
        const int rc = run_batch_chunk(b, n);
        if (rc) return rc;
        nsteps -= n;
    }
    return ADSG_OK;
}

int run_batch_chunk(adsg_batch* b, int nsteps) {
    const auto key = std::make_pair(nsteps, b->phase);
    auto it = b->graphs.find(key);
    hipGraphExec_t ex = nullptr;
    if (it != b->graphs.end()) {
        ex = it->second;
    } else {
        const int phase0 = b->phase;
        hipError_t e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return bhip(b, e, "batch: hipStreamBeginCapture");
        int rc = ADSG_OK;
        for (int s = 0; s < nsteps && rc == ADSG_OK; ++s) rc = batch_step(b);
        hipGraph_t g = nullptr;
        e = hipStreamEndCapture(b->stream, &g);
        b->phase = phase0;  // capture only recorded the steps
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return bhip(b, e, "batch: hipStreamEndCapture");
        e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return bhip(b, e, "batch: hipGraphInstantiate");
        b->graphs[key] = ex;
    }
    hipError_t e = hipGraphLaunch(ex, b->stream);
    if (e != hipSuccess) return bhip(b, e, "batch: hipGraphLaunch");
    b->phase ^= (nsteps & 1);
    return ADSG_OK;
}

bool valid_problem(const adsg_batch* b, int i) { return b && i >= 0 && i < static_cast<int>(b->sims.size()); }

}  // namespace

extern "C" {

const char* adsg_batch_last_error(const adsg_batch* b) { return b ? b->err.c_str() : g_batch_error.c_str(); }

int adsg_batch_create(const adsg_config* cfgs, int count, adsg_batch** out) {
    if (!cfgs || !out || count <= 0) {
        g_batch_error = "batch_create: bad argument";
        return ADSG_EINVAL;
    }
    *out = nullptr;
    auto* b = new adsg_batch;
    b->precision = cfgs[0].precision;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        g_batch_error = "batch_create: hipStreamCreate failed (no HIP device?)";
        delete b;
        return ADSG_EHIP;
    }
    for (int i = 0; i < count; ++i) {
        const adsg_config& c = cfgs[i];
        if (c.dim != 2 || c.precision != b->precision || c.x.p != c.y.p || c.x.p < 1 || c.x.p > 3) {
            g_batch_error = "batch_create: problem " + std::to_string(i) +
                            ": batches hold 2D problems of degree 1..3 (same in x and y) and one precision";
            adsg_batch_destroy(b);
            return ADSG_EINVAL;
        }
    }
    // the problems' host set-up on worker threads, each staging its tables; then one upload
    std::vector<adsg::stager> st(count);
    std::vector<adsg_sim*> hs(count, nullptr);
    std::vector<int> rcs(count, ADSG_OK);
    std::vector<std::string> msgs(count);
    for_each_problem(count, [&](int i) {
        rcs[i] = adsg::create_sim(&cfgs[i], b->stream, false, &hs[i], &st[i]);
        if (rcs[i]) msgs[i] = adsg_last_error(nullptr);  // this thread's create error
    });
    int bad = -1;
    for (int i = 0; i < count && bad < 0; ++i)
        if (rcs[i] || hs[i]->general) bad = i;
    for (int i = 0; i < count; ++i)
        if (hs[i]) {
            b->sims.push_back(hs[i]);
            b->off.push_back(b->total);
            b->total += hs[i]->N;
        }
    if (bad >= 0) {
        g_batch_error = "batch_create: problem " + std::to_string(bad) + ": " +
                        (rcs[bad] ? msgs[bad]
                                  : std::string("batches take p+1 Gauss points and simple knots (quad_order 0 or p+1, "
                                                "repeated_nodes 0)"));
        const int rc = rcs[bad] ? rcs[bad] : ADSG_EINVAL;
        adsg_batch_destroy(b);
        return rc;
    }
    adsg::stager tables;
    for (auto& x : st) tables.take(x);
    if (tables.commit(&b->arena[0]) != hipSuccess) {
        g_batch_error = "batch_create: uploading the problems' tables failed";
        adsg_batch_destroy(b);
        return ADSG_ENOMEM;
    }
    bool ok = true;
    for (auto& a : b->u) ok = ok && hipMalloc(reinterpret_cast<void**>(&a), b->total * sizeof(double)) == hipSuccess;
    if (b->precision == ADSG_FP32) {
        for (auto& a : b->sh) ok = ok && hipMalloc(reinterpret_cast<void**>(&a), b->total * sizeof(float)) == hipSuccess;
        ok = ok && hipMalloc(reinterpret_cast<void**>(&b->buf), b->total * sizeof(float)) == hipSuccess;
    }
    if (!ok) {
        g_batch_error = "batch_create: device allocation of the arenas failed";
        adsg_batch_destroy(b);
        return ADSG_ENOMEM;
    }
    for (auto& a : b->u) (void)hipMemsetAsync(a, 0, b->total * sizeof(double), b->stream);
    for (auto& a : b->sh)
        if (a) (void)hipMemsetAsync(a, 0, b->total * sizeof(float), b->stream);
    if (b->buf) (void)hipMemsetAsync(b->buf, 0, b->total * sizeof(float), b->stream);
    (void)hipStreamSynchronize(b->stream);
    b->forms.assign(count, adsg_rhs_form{});
    for (auto& f : b->forms) {
        f.mass = 1.0;
        f.grad[0] = f.grad[1] = f.grad[2] = 1.0;
    }
    *out = b;
    return ADSG_OK;
}

void adsg_batch_destroy(adsg_batch* b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_batch_graphs(b);
    free_groups(b);
    for (adsg_sim* h : b->sims) adsg_destroy(h);
    for (void* a : b->arena)
        if (a) (void)hipFree(a);
    for (auto& a : b->u)
        if (a) (void)hipFree(a);
    for (auto& a : b->sh)
        if (a) (void)hipFree(a);
    if (b->buf) (void)hipFree(b->buf);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int adsg_batch_count(const adsg_batch* b) { return b ? static_cast<int>(b->sims.size()) : ADSG_EINVAL; }

long adsg_batch_size(const adsg_batch* b, int problem) {
    return valid_problem(b, problem) ? b->sims[problem]->N : ADSG_EINVAL;
}

int adsg_batch_dofs(const adsg_batch* b, int problem, int dir) {
    return valid_problem(b, problem) ? adsg_dofs(b->sims[problem], dir) : ADSG_EINVAL;
}

int adsg_batch_prepare(adsg_batch* b) {
    if (!b) return ADSG_EINVAL;
    if (b->prepared) return ADSG_OK;  // the factors live in the batch's arena
    // every problem's factors on worker threads, each staged; then one upload
    const int count = static_cast<int>(b->sims.size());
    std::vector<adsg::stager> st(count);
    std::vector<int> rcs(count, ADSG_OK);
    for_each_problem(count, [&](int i) {
        adsg_sim* h = b->sims[i];
        h->stage = &st[i];
        rcs[i] = adsg_prepare_matrices(h);
        h->stage = nullptr;
    });
    adsg::stager factors;
    for (auto& x : st) factors.take(x);
    for (size_t i = 0; i < b->sims.size(); ++i) {
        adsg_sim* h = b->sims[i];
        const int rc = rcs[i];
        if (rc) return bfail(b, rc, "batch_prepare: problem " + std::to_string(i) + ": " + h->err);
        for (int d = 0; d < 2; ++d) {
            const auto& s = h->dirs[d].slots[0];
            if (s.lf.pivoted || s.lf.kl != h->dirs[d].tab.p || s.lf.ku_eff != h->dirs[d].tab.p)
                return bfail(b, ADSG_EINVAL, "batch_prepare: problem " + std::to_string(i) +
                                                 ": batched sweeps need unpivoted degree-p band factors");
        }
    }
    hipError_t e = factors.commit(&b->arena[1]);
    if (e != hipSuccess) return bhip(b, e, "batch_prepare: uploading the factors");
    int rc = build(b);
    if (rc) return rc;
    b->prepared = true;
    return ADSG_OK;
}

int adsg_batch_set_forms(adsg_batch* b, const adsg_rhs_form* forms) {
    if (!b || !forms) return bfail(b, ADSG_EINVAL, "batch_set_forms: bad argument");
    for (size_t i = 0; i < b->sims.size(); ++i)
        if (forms[i].use_kfield || forms[i].use_source)
            return bfail(b, ADSG_EINVAL, "batch_set_forms: batched forms take no k field or source");
    b->forms.assign(forms, forms + b->sims.size());
    return b->prepared ? build(b) : ADSG_OK;
}

int adsg_batch_set_vector(adsg_batch* b, int problem, int which, const double* host) {
    if (!valid_problem(b, problem) || !host || (which != ADSG_U && which != ADSG_U_PREV))
        return bfail(b, ADSG_EINVAL, "batch_set_vector: bad argument (U or U_PREV of a problem)");
    const int a = which == ADSG_U ? b->phase : 1 - b->phase;
    const long o = b->off[problem], n = b->sims[problem]->N;
    hipError_t e = hipStreamSynchronize(b->stream);
    if (e == hipSuccess) e = hipMemcpy(b->u[a] + o, host, n * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess && b->precision == ADSG_FP32) {
        e = adsg::launch_convert_d2f(b->u[a] + o, b->sh[a] + o, n, b->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    }
    return e == hipSuccess ? ADSG_OK : bhip(b, e, "batch_set_vector");
}

int adsg_batch_get_vector(adsg_batch* b, int problem, int which, double* host) {
    if (!valid_problem(b, problem) || !host || (which != ADSG_U && which != ADSG_U_PREV))
        return bfail(b, ADSG_EINVAL, "batch_get_vector: bad argument (U or U_PREV of a problem)");
    const int a = which == ADSG_U ? b->phase : 1 - b->phase;
    hipError_t e = hipStreamSynchronize(b->stream);
    if (e == hipSuccess)
        e = hipMemcpy(host, b->u[a] + b->off[problem], b->sims[problem]->N * sizeof(double),
                      hipMemcpyDeviceToHost);
    return e == hipSuccess ? ADSG_OK : bhip(b, e, "batch_get_vector");
}

int adsg_batch_step(adsg_batch* b, int nsteps) {
    if (!b || nsteps < 0) return bfail(b, ADSG_EINVAL, "batch_step: bad argument");
    if (!b->prepared) return bfail(b, ADSG_ESTATE, "batch_step: batch_prepare first");
    if (nsteps == 0) return ADSG_OK;
    return run_batch(b, nsteps);
}

int adsg_batch_synchronize(adsg_batch* b) {
    if (!b) return ADSG_EINVAL;
    hipError_t e = hipStreamSynchronize(b->stream);
    return e == hipSuccess ? ADSG_OK : bhip(b, e, "batch_synchronize");
}

int adsg_batch_set_tuning(adsg_batch* b, int key, int value) {
    if (!b || (key != ADSG_TUNE_BATCH_TILE && key != ADSG_TUNE_SWEEP_VARIANT) || value < 0 || value > 1)
        return bfail(b, ADSG_EINVAL, "batch_set_tuning: unknown key or value");
    if (key == ADSG_TUNE_SWEEP_VARIANT) {
        if (b->sweep != value) drop_batch_graphs(b);
        b->sweep = value;
        return ADSG_OK;
    }
    if (b->shape == value) return ADSG_OK;
    b->shape = value;
    return b->prepared ? build(b) : ADSG_OK;
}

int adsg_batch_get_tuning(const adsg_batch* b, int key) {
    if (!b) return ADSG_EINVAL;
    if (key == ADSG_TUNE_BATCH_TILE) return b->shape;
    if (key == ADSG_TUNE_SWEEP_VARIANT) return b->sweep;
    return ADSG_EINVAL;
}

int adsg_batch_autotune(adsg_batch* b, double* ms_out) {
    if (!b) return ADSG_EINVAL;
    if (!b->prepared) return bfail(b, ADSG_ESTATE, "batch_autotune: batch_prepare first");
    // keep the state: copies of both arenas (and shadows); eager steps per tile shape
    const size_t nd = b->total * sizeof(double), nf = b->total * sizeof(float);
    std::vector<void*> saved;
    std::vector<std::pair<void*, size_t>> what = {{b->u[0], nd}, {b->u[1], nd}};
    if (b->precision == ADSG_FP32) {
        what.push_back({b->sh[0], nf});
        what.push_back({b->sh[1], nf});
    }
    for (auto& w : what) {
        void* p = nullptr;
        if (hipMalloc(&p, w.second) != hipSuccess) {
            for (void* q : saved) (void)hipFree(q);
            return bfail(b, ADSG_ENOMEM, "batch_autotune: no memory for the state copy");
        }
        (void)hipMemcpyAsync(p, w.first, w.second, hipMemcpyDeviceToDevice, b->stream);
        saved.push_back(p);
    }
    const int phase0 = b->phase;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    // every (tile shape, line grouping) pair: 16 x 16 or 8 x 8 element tiles; lines either
    // segmented (64 lines per workgroup of nine waves) or one lane per line (64 lines per
    // one-wave workgroup)
    int best = b->shape, best_sweep = b->sweep, rc = ADSG_OK;
    float best_ms = 1e30f;
    for (int shape = 0; shape <= 1 && rc == ADSG_OK; ++shape) {
        float shape_ms = 1e30f;
        for (int sweep = 0; sweep <= 1 && rc == ADSG_OK; ++sweep) {
            b->shape = shape;
            b->sweep = sweep;
            if ((rc = build(b))) break;
            rc = batch_step(b);  // code load
            (void)hipEventRecord(e0, b->stream);
            for (int s = 0; s < 2 && rc == ADSG_OK; ++s) rc = batch_step(b);
            (void)hipEventRecord(e1, b->stream);
            (void)hipEventSynchronize(e1);
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            shape_ms = ms < shape_ms ? ms : shape_ms;
            if (ms < best_ms) {
                best_ms = ms;
                best = shape;
                best_sweep = sweep;
            }
        }
        if (ms_out) ms_out[shape] = shape_ms / 2;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    for (size_t i = 0; i < what.size(); ++i)
        (void)hipMemcpyAsync(what[i].first, saved[i], what[i].second, hipMemcpyDeviceToDevice, b->stream);
    (void)hipStreamSynchronize(b->stream);
    for (void* q : saved) (void)hipFree(q);
    b->phase = phase0;
    b->shape = best;
    b->sweep = best_sweep;
    drop_batch_graphs(b);
    int rc2 = build(b);
    return rc ? rc : rc2;
}

}  // extern "C"
