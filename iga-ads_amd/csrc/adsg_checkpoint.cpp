// Checkpoint and resume of a handle's state (no counterpart in the reference, which writes
// text coefficient dumps, implicit.hpp:121-127, and never reads them back into a run).
//
// File layout (little-endian, as the host writes it):
//   header (128 bytes)
//     char     magic[8]      "ADSGSTAT"
//     uint32   version       1
//     uint32   dim           2 or 3
//     int32    p[3], elements[3], dofs[3], quad_order[3], repeated_nodes[3]
//     int32    precision     ADSG_FP64 / ADSG_FP32 of the handle that wrote it
//     int32    reserved
//     int64    step          the caller's step index
//     double   t             the caller's time
//     int64    n             coefficients per vector
//     uint64   checksum      FNV-1a 64 of the payload bytes
//     padding to 128 bytes
//   payload: u then u_prev, n FP64 values each (the state vectors are FP64 in both
//   precisions).
// A load checks the magic, version, the configuration against the handle's and the checksum
// before it touches the handle; a mismatch returns ADSG_EINVAL and leaves the state as it was.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "adsg_sim.hpp"

namespace {

struct ckpt_header {
    char magic[8];
    uint32_t version;
    uint32_t dim;
    int32_t p[3], elements[3], dofs[3], quad_order[3], repeated_nodes[3];
    int32_t precision;
    int32_t reserved;
    int64_t step;
    double t;
    int64_t n;
    uint64_t checksum;
};
static_assert(sizeof(ckpt_header) <= 128, "header fits its 128 bytes");
constexpr char kMagic[8] = {'A', 'D', 'S', 'G', 'S', 'T', 'A', 'T'};
constexpr uint32_t kVersion = 1;

uint64_t fnv1a(const void* data, size_t bytes, uint64_t h = 1469598103934665603ULL) {
    const unsigned char* p = static_cast<const unsigned char*>(data);
    for (size_t i = 0; i < bytes; ++i) {
        h ^= p[i];
        h *= 1099511628211ULL;
    }
    return h;
}

int fail(adsg_sim* h, int code, const std::string& msg) {
    if (h) h->err = msg;
    return code;
}

ckpt_header header_of(const adsg_sim* h) {
    ckpt_header c{};
    std::memcpy(c.magic, kMagic, 8);
    c.version = kVersion;
    c.dim = static_cast<uint32_t>(h->dim);
    for (int d = 0; d < 3; ++d) {
        const bool on = d < h->dim;
        c.p[d] = on ? h->dirs[d].tab.p : 0;
        c.elements[d] = on ? h->dirs[d].tab.elements : 0;
        c.dofs[d] = on ? h->dirs[d].tab.dofs : 0;
        c.quad_order[d] = on ? h->dirs[d].tab.nq : 0;
        c.repeated_nodes[d] = on ? h->dirs[d].tab.repeated_nodes : 0;
    }
    c.precision = h->precision;
    c.n = h->N;
    return c;
}

}  // namespace

extern "C" {

int adsg_save_state(adsg_sim* h, const char* path, long step, double t) {
    if (!h || !path) return fail(h, ADSG_EINVAL, "save_state: bad argument");
    if (!h->vec[ADSG_U] || !h->vec[ADSG_U_PREV]) return fail(h, ADSG_ESTATE, "save_state: no state vectors");
    std::vector<double> data(2 * static_cast<size_t>(h->N));
    for (int k = 0; k < 2; ++k) {
        const int rc = adsg_get_vector(h, k == 0 ? ADSG_U : ADSG_U_PREV, data.data() + k * h->N);
        if (rc) return rc;
    }
    ckpt_header c = header_of(h);
    c.step = step;
    c.t = t;
    c.checksum = fnv1a(data.data(), data.size() * sizeof(double));
    unsigned char head[128] = {};
    std::memcpy(head, &c, sizeof(c));
    const std::string tmp = std::string(path) + ".tmp";
    FILE* f = std::fopen(tmp.c_str(), "wb");
    if (!f) return fail(h, ADSG_EINVAL, "save_state: cannot open " + tmp + " for writing");
    const bool ok = std::fwrite(head, 1, sizeof(head), f) == sizeof(head) &&
                    std::fwrite(data.data(), sizeof(double), data.size(), f) == data.size();
    const bool closed = std::fclose(f) == 0;
    if (!ok || !closed || std::rename(tmp.c_str(), path) != 0) {
        std::remove(tmp.c_str());
        return fail(h, ADSG_EINVAL, std::string("save_state: writing ") + path + " failed");
    }
    return ADSG_OK;
}

int adsg_load_state(adsg_sim* h, const char* path, long* step, double* t) {
    if (!h || !path) return fail(h, ADSG_EINVAL, "load_state: bad argument");
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(h, ADSG_EINVAL, std::string("load_state: cannot open ") + path);
    unsigned char head[128];
    ckpt_header c{};
    const bool got_head = std::fread(head, 1, sizeof(head), f) == sizeof(head);
    if (got_head) std::memcpy(&c, head, sizeof(c));
    std::string why;
    const ckpt_header mine = header_of(h);
    if (!got_head || std::memcmp(c.magic, kMagic, 8) != 0) {
        why = "not a state file";
    } else if (c.version != kVersion) {
        why = "unknown version " + std::to_string(c.version);
    } else if (c.dim != mine.dim || c.n != mine.n ||
               std::memcmp(c.p, mine.p, sizeof(c.p)) != 0 || std::memcmp(c.elements, mine.elements, sizeof(c.elements)) != 0 ||
               std::memcmp(c.dofs, mine.dofs, sizeof(c.dofs)) != 0 ||
               std::memcmp(c.quad_order, mine.quad_order, sizeof(c.quad_order)) != 0 ||
               std::memcmp(c.repeated_nodes, mine.repeated_nodes, sizeof(c.repeated_nodes)) != 0) {
        why = "written for another configuration";
    }
    std::vector<double> data;
    if (why.empty()) {
        data.resize(2 * static_cast<size_t>(c.n));
        if (std::fread(data.data(), sizeof(double), data.size(), f) != data.size())
            why = "truncated";
        else if (std::fgetc(f) != EOF)
            why = "longer than its header says";
        else if (fnv1a(data.data(), data.size() * sizeof(double)) != c.checksum)
            why = "checksum mismatch (corrupted)";
    }
    std::fclose(f);
    if (!why.empty()) return fail(h, ADSG_EINVAL, std::string("load_state: ") + path + ": " + why);
    for (int k = 0; k < 2; ++k) {
        const int rc = adsg_set_vector(h, k == 0 ? ADSG_U : ADSG_U_PREV, data.data() + k * c.n);
        if (rc) return rc;
    }
    const int rc = adsg_synchronize(h);
    if (rc) return rc;
    if (step) *step = static_cast<long>(c.step);
    if (t) *t = c.t;
    return ADSG_OK;
}

}  // extern "C"
