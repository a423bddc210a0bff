// Batches of independent 2D problems advanced together (declared in include/adsgpu.h).
//
// Every problem keeps its own 1D tables and FP64-computed factors (an adsg_sim without state
// vectors), while the state of all problems lives in shared arenas, problem after problem.
// One step is a fixed sequence of grouped launches over all problems: zero fill of the
// arena, per degree p and parity colour one element-kernel launch whose blocks cover the
// tiles of every problem of that degree, per degree one x-sweep and one y-sweep launch over
// all their lines, and one elementwise update of the whole arena. The step sequences are
// replayed from captured HIP graphs of at most 50 steps. The problems' set-up runs on the
// device (adsg_batch_setup.hip): create tabulates every direction's basis in one launch,
// prepare assembles the Gram bands, factors them and packs the sweep rows in two more; the
// host only lays out the arenas and uploads their descriptors (one copy per phase).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstring>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "adsg_batch_setup.hpp"
#include "adsg_sim.hpp"

namespace {

thread_local std::string g_batch_error;

// the launches of one step for one arena parity, over every problem of every degree (device
// arrays of items, one item per problem)
struct batch_group {
    int nitems = 0;
    void* rhs_items = nullptr;  // rhs2_item<T>[nitems], the problems of degree 1, then 2, then 3
    int deg_first[4] = {0, 0, 0, 0};  // first item of degree p
    int deg_count[4] = {0, 0, 0, 0};  // items of degree p
    int deg_blocks[4][4] = {};        // element-kernel blocks of degree p in colour c
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
    void* items = nullptr;  // the launch items of every group, one allocation
    void* packs = nullptr;  // packed sweep rows of every problem and direction (step precision)
    std::map<std::pair<int, int>, hipGraphExec_t> graphs;  // (nsteps, phase at start)
    // every problem's tables (create) and factors (prepare), each set in one device arena,
    // laid out by the descriptors of its directions (problem i: 2i x, 2i + 1 y)
    void* arena[2] = {nullptr, nullptr};
    std::vector<adsg::bs_dir> bs;
    std::string err;
};

namespace {

int bfail(adsg_batch* b, int code, const std::string& msg) {
    if (b) b->err = msg;
    return code;
}

int bhip(adsg_batch* b, hipError_t e, const char* what) {
    return bfail(b, ADSG_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

void free_groups(adsg_batch* b) {
    for (auto& gs : b->groups) gs.clear();
    if (b->items) (void)hipFree(b->items);
    b->items = nullptr;
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

// Host bytes of the launch items of every group, copied to the device in one allocation:
// add() appends an item array (256-byte aligned) and returns its byte offset.
struct item_stage {
    std::vector<char> bytes;
    template <typename T>
    size_t add(const std::vector<T>& items) {
        const size_t o = (bytes.size() + 255) / 256 * 256;
        bytes.resize(o + sizeof(T) * items.size());
        if (!items.empty()) std::memcpy(bytes.data() + o, items.data(), sizeof(T) * items.size());
        return o;
    }
};

// launch items of every problem for both arena parities (T: the step precision)
template <typename T>
int build_groups(adsg_batch* b) {
    free_groups(b);
    drop_batch_graphs(b);
    // the packed sweep rows come from prepare (every direction segmented, or none)
    const bool all_seg = b->packs != nullptr;
    const int P = static_cast<int>(b->sims.size());
    item_stage stage;
    std::vector<std::array<size_t, 5>> offs;  // per group, in the order of b->groups[0], [1]
    for (int ph = 0; ph < 2; ++ph) {
        std::vector<adsg::rhs2_item<T>> ri;
        std::vector<adsg::sweep_item<T>> sx, sy;
        std::vector<adsg::seg_item<T>> gx, gy;
        batch_group g;
        std::vector<adsg::rhs2_item<T>> rdeg[4];
        for (int i = 0; i < P; ++i) {
            const adsg_sim* h = b->sims[i];
            const int p = h->dirs[0].tab.p;
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
            it.prm.store = true;  // no zero fill: the first colour to reach a DOF stores
            int count[4];
            adsg::rhs2_batch_tiles(b->shape, it.X.elements, it.Y.elements, it.ntx2, count);
            for (int c = 0; c < 4; ++c) {
                it.first[c] = g.deg_blocks[p][c];
                g.deg_blocks[p][c] += count[c];
            }
            rdeg[p].push_back(it);
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
                const adsg::bs_dir& sgx = b->bs[2 * i];
                const adsg::bs_dir& sgy = b->bs[2 * i + 1];
                a.u = c.u = target;
                a.g.n_lines = ny;  // x line l starts at l * nx
                a.g.A = 1;
                a.g.B = nx;
                a.g.S = 1;
                a.pack = static_cast<const T*>(sgx.pack);
                a.n = static_cast<int>(nx);
                a.m = sgx.m;
                a.nseg = sgx.nseg;
                a.p = p;
                a.first = g.gx_blocks;
                g.gx_blocks += static_cast<int>((ny + 63) / 64);
                c.g = yi.g;
                if constexpr (sizeof(T) == sizeof(float)) {  // the y sweep does the FP32 step's update
                    c.x_prev = b->u[ph] + b->off[i];
                    c.x_new = b->u[1 - ph] + b->off[i];
                    c.shadow = b->sh[1 - ph] + b->off[i];
                }
                c.pack = static_cast<const T*>(sgy.pack);
                c.n = static_cast<int>(ny);
                c.m = sgy.m;
                c.nseg = sgy.nseg;
                c.p = p;
                c.first = g.gy_blocks;
                g.gy_blocks += static_cast<int>((nx + 63) / 64);
                gx.push_back(a);
                gy.push_back(c);
            }
        }
        // the element items by degree
        for (int p = 1; p <= 3; ++p) {
            g.deg_first[p] = static_cast<int>(ri.size());
            g.deg_count[p] = static_cast<int>(rdeg[p].size());
            ri.insert(ri.end(), rdeg[p].begin(), rdeg[p].end());
        }
        if (ri.empty()) continue;
        g.nitems = static_cast<int>(ri.size());
        g.seg = all_seg;
        // byte offsets of the group's item arrays, turned into device pointers below
        offs.push_back({stage.add(ri), stage.add(sx), stage.add(sy), all_seg ? stage.add(gx) : 0,
                        all_seg ? stage.add(gy) : 0});
        b->groups[ph].push_back(g);
    }
    if (stage.bytes.empty()) return ADSG_OK;
    if (hipMalloc(&b->items, stage.bytes.size()) != hipSuccess)
        return bfail(b, ADSG_ENOMEM, "batch: no memory for the launch items");
    hipError_t e = hipMemcpy(b->items, stage.bytes.data(), stage.bytes.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) return bhip(b, e, "batch: uploading launch items");
    char* base = static_cast<char*>(b->items);
    size_t k = 0;
    for (auto& gs : b->groups)
        for (batch_group& g : gs) {
            const auto& o = offs[k++];
            g.rhs_items = base + o[0];
            g.sx_items = base + o[1];
            g.sy_items = base + o[2];
            g.gx_items = g.seg ? base + o[3] : nullptr;
            g.gy_items = g.seg ? base + o[4] : nullptr;
        }
    return ADSG_OK;
}

int build(adsg_batch* b) {
    return b->precision == ADSG_FP32 ? build_groups<float>(b) : build_groups<double>(b);
}

// one step of every problem, starting at arena parity b->phase: the element kernels of the
// problems of each degree in one launch per colour (the first colour to reach a DOF stores:
// no zero fill), the x lines of every problem (every degree) in one launch, the y lines in
// one, and (FP32) the update: done by the segmented y sweeps as they store their lines,
// else a pass of its own
template <typename T>
int batch_step_impl(adsg_batch* b) {
    const int ph = b->phase;
    hipError_t e = hipSuccess;
    for (const batch_group& g : b->groups[ph]) {
        const auto* items = static_cast<const adsg::rhs2_item<T>*>(g.rhs_items);
        for (int p = 1; p <= 3 && e == hipSuccess; ++p)
            for (int c = 0; c < 4 && e == hipSuccess; ++c)
                e = adsg::launch_rhs2_batch<T>(p, b->shape, items + g.deg_first[p], g.deg_count[p], c,
                                               g.deg_blocks[p][c], b->stream);
        if (e != hipSuccess) return bhip(b, e, "batch: element kernels");
        if (g.seg && b->sweep == 0)
            e = adsg::launch_sweep_seg_batch<T>(static_cast<const adsg::seg_item<T>*>(g.gx_items), g.nitems,
                                                g.gx_blocks, b->stream);
        else
            e = adsg::launch_sweep_batch<T>(static_cast<const adsg::sweep_item<T>*>(g.sx_items), g.nitems,
                                            g.sx_blocks, b->stream);
        if (e != hipSuccess) return bhip(b, e, "batch: x sweeps");
        if (g.seg && b->sweep == 0)
            e = adsg::launch_sweep_seg_batch<T>(static_cast<const adsg::seg_item<T>*>(g.gy_items), g.nitems,
                                                g.gy_blocks, b->stream);
        else
            e = adsg::launch_sweep_batch<T>(static_cast<const adsg::sweep_item<T>*>(g.sy_items), g.nitems,
                                            g.sy_blocks, b->stream);
        if (e != hipSuccess) return bhip(b, e, "batch: y sweeps");
    }
    if constexpr (sizeof(T) == sizeof(float)) {
        // u = u_prev + increment in FP64, with its FP32 copy (no boundary rows: natural BC)
        bool updated = true;  // by the segmented y sweeps (seg_item::x_new), in every group
        for (const batch_group& g : b->groups[ph]) updated = updated && g.seg && b->sweep == 0;
        if (!updated) {
            e = adsg::launch_apply_increment(b->u[ph], b->buf, b->u[1 - ph], b->sh[1 - ph], 2,
                                             static_cast<int>(b->total), 1, 1, false, b->stream);
            if (e != hipSuccess) return bhip(b, e, "batch: update");
        }
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

// A problem of the batch: the handle without state vectors, tables or factors of its own
// (they are laid out in the batch's arenas by the device set-up below).
adsg_sim* batch_sim(const adsg_config& c, hipStream_t stream, int device) {
    auto* h = new adsg_sim;
    h->dim = 2;
    h->precision = c.precision;
    h->step_count = c.step_count;
    h->dt = c.dt;
    h->device = device;
    h->stream = stream;
    h->owns_stream = false;
    h->in_arena = true;
    const adsg_dim_config* dc[2] = {&c.x, &c.y};
    h->N = 1;
    for (int d = 0; d < 2; ++d) {
        adsg::direction& t = h->dirs[d].tab;
        t.p = dc[d]->p;
        t.elements = dc[d]->elements;
        t.dofs = t.elements + t.p;  // clamped knots, simple interior knots
        t.nq = t.p + 1;
        t.a = dc[d]->a;
        t.b = dc[d]->b;
        h->N *= t.dofs;
        adsg::dev_slot s0;
        s0.set = true;
        s0.kl = s0.ku = t.p;
        h->dirs[d].slots.push_back(std::move(s0));
    }
    return h;
}

// byte layout of an arena: take() returns the 256-aligned offset of the next `bytes`
struct arena_layout {
    size_t size = 0;
    size_t take(size_t bytes) {
        const size_t o = size;
        size += (bytes + 255) / 256 * 256;
        return o;
    }
};

// Gauss-Legendre rules of 2, 3 and 4 points (p = 1..3), nodes then weights, at
// kGaussOff[p] in one array of 18 values
constexpr int kGaussOff[4] = {0, 0, 4, 10};

// create: arena 0 = [Gauss rules | descriptors | every direction's val, der, wj (T)], one
// copy of the head, one launch for the tables
template <typename T>
int batch_tables(adsg_batch* b) {
    const int count = static_cast<int>(2 * b->sims.size());
    b->bs.assign(count, adsg::bs_dir{});
    arena_layout L;
    const size_t g_off = L.take(18 * sizeof(double));
    const size_t d_off = L.take(count * sizeof(adsg::bs_dir));
    std::vector<std::array<size_t, 3>> to(count);
    for (int k = 0; k < count; ++k) {
        const adsg::direction& t = b->sims[k / 2]->dirs[k % 2].tab;
        const size_t m = t.p + 1;
        to[k] = {L.take(sizeof(T) * t.elements * m * m), L.take(sizeof(T) * t.elements * m * m),
                 L.take(sizeof(T) * t.elements * m)};
    }
    if (hipMalloc(&b->arena[0], L.size) != hipSuccess)
        return bfail(b, ADSG_ENOMEM, "batch_create: no memory for the problems' tables");
    char* base = static_cast<char*>(b->arena[0]);
    std::vector<char> head(d_off + count * sizeof(adsg::bs_dir), 0);
    double* g = reinterpret_cast<double*>(head.data() + g_off);
    for (int p = 1; p <= 3; ++p) adsg::gauss_legendre(p + 1, g + kGaussOff[p], g + kGaussOff[p] + p + 1);
    const double* gdev = reinterpret_cast<const double*>(base + g_off);
    for (int k = 0; k < count; ++k) {
        adsg_sim* h = b->sims[k / 2];
        adsg::dev_dir& dd = h->dirs[k % 2];
        adsg::bs_dir& d = b->bs[k];
        d.p = dd.tab.p;
        d.ne = dd.tab.elements;
        d.n = dd.tab.dofs;
        d.a = dd.tab.a;
        d.b = dd.tab.b;
        d.gx = gdev + kGaussOff[d.p];
        d.gw = gdev + kGaussOff[d.p] + d.p + 1;
        d.val = dd.val = base + to[k][0];
        d.der = dd.der = base + to[k][1];
        d.wj = dd.wj = base + to[k][2];
    }
    std::memcpy(head.data() + d_off, b->bs.data(), count * sizeof(adsg::bs_dir));
    hipError_t e = hipMemcpyAsync(base, head.data(), head.size(), hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess)
        e = adsg::launch_batch_tables<T>(reinterpret_cast<const adsg::bs_dir*>(base + d_off), count, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    return e == hipSuccess ? ADSG_OK : bhip(b, e, "batch_create: tabulating the bases");
}

// prepare: arena 1 = [descriptors | info | every direction's lower, upper, rdiag (T), poff];
// the packed sweep rows (when every direction splits into segments) in b->packs; the FP64
// Gram bands in a work array freed at the end
template <typename T>
int batch_factors(adsg_batch* b) {
    const int count = static_cast<int>(b->bs.size());
    for (void** a : {&b->arena[1], &b->packs})  // a failed earlier prepare
        if (*a) {
            (void)hipFree(*a);
            *a = nullptr;
        }
    bool all_seg = true;
    for (auto& d : b->bs) {
        int m = 0, nseg = 0, mseg = 0;
        if (!adsg::segment_shape(d.n, d.p, 0, adsg::kBatchSegments, m, nseg, mseg)) all_seg = false;
        d.m = m;
        d.nseg = nseg;
        d.mseg = mseg;
    }
    arena_layout L, W, Pk;
    const size_t d_off = L.take(count * sizeof(adsg::bs_dir));
    const size_t i_off = L.take(count * sizeof(int));
    std::vector<std::array<size_t, 6>> fo(count);
    for (int k = 0; k < count; ++k) {
        const adsg::bs_dir& d = b->bs[k];
        fo[k] = {L.take(sizeof(T) * d.n * d.p), L.take(sizeof(T) * d.n * d.p), L.take(sizeof(T) * d.n),
                 L.take(sizeof(int) * d.n), W.take(sizeof(double) * d.n * (3 * d.p + 1)),
                 all_seg ? Pk.take(sizeof(T) * d.nseg * d.mseg * (4 * d.p + 1)) : 0};
    }
    void* work = nullptr;
    if (hipMalloc(&b->arena[1], L.size) != hipSuccess || hipMalloc(&work, W.size) != hipSuccess ||
        (all_seg && hipMalloc(&b->packs, Pk.size) != hipSuccess)) {
        if (work) (void)hipFree(work);
        return bfail(b, ADSG_ENOMEM, "batch_prepare: no memory for the factors");
    }
    char* base = static_cast<char*>(b->arena[1]);
    for (int k = 0; k < count; ++k) {
        adsg::bs_dir& d = b->bs[k];
        d.lower = base + fo[k][0];
        d.upper = base + fo[k][1];
        d.rdiag = base + fo[k][2];
        d.poff = reinterpret_cast<int*>(base + fo[k][3]);
        d.ab = reinterpret_cast<double*>(static_cast<char*>(work) + fo[k][4]);
        d.info = reinterpret_cast<int*>(base + i_off) + k;
        if (!all_seg) d.nseg = 0;
        d.pack = all_seg ? static_cast<char*>(b->packs) + fo[k][5] : nullptr;
    }
    std::vector<int> info(count, 0);
    hipError_t e = hipMemcpyAsync(base + d_off, b->bs.data(), count * sizeof(adsg::bs_dir), hipMemcpyHostToDevice,
                                  b->stream);
    if (e == hipSuccess)
        e = adsg::launch_batch_factors<T>(reinterpret_cast<const adsg::bs_dir*>(base + d_off), count, b->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(info.data(), base + i_off, count * sizeof(int), hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    (void)hipFree(work);
    for (auto& d : b->bs) d.ab = nullptr;
    if (e != hipSuccess) return bhip(b, e, "batch_prepare: factoring");
    for (int k = 0; k < count; ++k) {
        if (info[k] == -1)
            return bfail(b, ADSG_EINVAL, "batch_prepare: problem " + std::to_string(k / 2) +
                                             ": batched sweeps need unpivoted degree-p band factors");
        if (info[k] > 0)
            return bfail(b, info[k], "batch_prepare: problem " + std::to_string(k / 2) + ": zero pivot at row " +
                                         std::to_string(info[k]));
    }
    for (int k = 0; k < count; ++k) {
        const adsg::bs_dir& d = b->bs[k];
        adsg_sim* h = b->sims[k / 2];
        adsg::dev_slot& s = h->dirs[k % 2].slots[0];
        s.lower = d.lower;
        s.upper = d.upper;
        s.rdiag = d.rdiag;
        s.poff = d.poff;
        s.lf.n = d.n;
        s.lf.kl = s.lf.ku = s.lf.ku_eff = d.p;
        s.lf.pivoted = false;
        s.factored = true;
        h->prepared = true;
    }
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
    int device = 0;
    (void)hipGetDevice(&device);
    for (int i = 0; i < count; ++i) {
        const adsg_config& c = cfgs[i];
        if (c.dim != 2 || c.precision != b->precision || c.x.p != c.y.p || c.x.p < 1 || c.x.p > 3) {
            g_batch_error = "batch_create: problem " + std::to_string(i) +
                            ": batches hold 2D problems of degree 1..3 (same in x and y) and one precision";
            adsg_batch_destroy(b);
            return ADSG_EINVAL;
        }
        for (const adsg_dim_config* dc : {&c.x, &c.y}) {
            if (dc->elements < 1 || !(dc->a < dc->b)) {
                g_batch_error = "batch_create: problem " + std::to_string(i) +
                                ": need at least one element and a < b in every direction";
                adsg_batch_destroy(b);
                return ADSG_EINVAL;
            }
            if ((dc->quad_order > 0 && dc->quad_order != dc->p + 1) || dc->repeated_nodes != 0) {
                g_batch_error = "batch_create: problem " + std::to_string(i) +
                                ": batches take p+1 Gauss points and simple knots (quad_order 0 or p+1, "
                                "repeated_nodes 0)";
                adsg_batch_destroy(b);
                return ADSG_EINVAL;
            }
        }
    }
    for (int i = 0; i < count; ++i) {
        b->sims.push_back(batch_sim(cfgs[i], b->stream, device));
        b->off.push_back(b->total);
        b->total += b->sims.back()->N;
    }
    // every problem's basis tables, tabulated on the device
    if (const int rc = b->precision == ADSG_FP32 ? batch_tables<float>(b) : batch_tables<double>(b)) {
        g_batch_error = b->err;
        adsg_batch_destroy(b);
        return rc;
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
    if (b->packs) (void)hipFree(b->packs);
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
    // every problem's Gram bands, their LU and the packed sweep rows, on the device
    int rc = b->precision == ADSG_FP32 ? batch_factors<float>(b) : batch_factors<double>(b);
    if (rc) return rc;
    rc = build(b);
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
