// CSV measurement ingest and output: the C ABI of include/traccc_amd/io.h.
//
// Reading follows traccc::io::csv::read_measurements
// (io/src/csv/read_measurements.cpp:20-99) and make_measurement_edm
// (io/src/csv/make_measurement_edm.cpp:19-71): columns by header name,
// local_key taken as one raw byte, its bits 1 / 2 selecting loc0 / loc1 (in
// that order, so the subspace lists the measured bound indices), then an
// optional stable sort by the traccc measurement ordering with the old ->
// new row map.
#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <numeric>
#include <string>
#include <vector>

#include "../common/errors.hpp"
#include "traccc_amd/io.h"

struct ckf_measurement_buffer {
    std::vector<float> pos, var, time;
    std::vector<uint32_t> dims, ident, new_index;
    std::vector<uint64_t> surface;
    std::vector<uint8_t> subspace;
};

namespace {

using tamd::ckf_error;
using tamd::guarded;

enum col : int {
    kMeasId, kGeoId, kLocalKey, kLocal0, kLocal1, kPhi, kTheta, kTime, kVar0, kVar1,
    kVarPhi, kVarTheta, kVarTime, kNCols
};
const char* const kColNames[kNCols] = {"measurement_id", "geometry_id", "local_key", "local0",
                                       "local1", "phi", "theta", "time", "var_local0",
                                       "var_local1", "var_phi", "var_theta", "var_time"};

std::vector<std::string> split(const std::string& line) {
    std::vector<std::string> out;
    size_t a = 0;
    for (;;) {
        const size_t b = line.find(',', a);
        out.push_back(line.substr(a, b == std::string::npos ? std::string::npos : b - a));
        if (b == std::string::npos) break;
        a = b + 1;
    }
    return out;
}

[[noreturn]] void bad(const std::string& path, size_t line_no, const std::string& what) {
    throw ckf_error(CKF_ERR_INVALID_ARGUMENT,
                    path + ":" + std::to_string(line_no) + ": " + what);
}

float to_float(const std::string& s, const std::string& path, size_t ln) {
    if (s.empty()) return 0.f;
    char* end = nullptr;
    errno = 0;
    const float v = std::strtof(s.c_str(), &end);
    if (end == s.c_str() || *end != '\0' || errno == ERANGE) bad(path, ln, "not a number: '" + s + "'");
    return v;
}

uint64_t to_u64(const std::string& s, const std::string& path, size_t ln) {
    if (s.empty()) return 0u;
    char* end = nullptr;
    errno = 0;
    const unsigned long long v = std::strtoull(s.c_str(), &end, 10);
    if (end == s.c_str() || *end != '\0' || errno == ERANGE) bad(path, ln, "not an integer: '" + s + "'");
    return static_cast<uint64_t>(v);
}

// traccc measurement ordering (measurement_collection.ipp:41-57)
bool meas_less(const ckf_measurement_buffer& B, uint32_t i, uint32_t j) {
    if (B.surface[i] != B.surface[j]) return B.surface[i] < B.surface[j];
    if (B.pos[2 * i] != B.pos[2 * j]) return B.pos[2 * i] < B.pos[2 * j];
    if (B.pos[2 * i + 1] != B.pos[2 * j + 1]) return B.pos[2 * i + 1] < B.pos[2 * j + 1];
    if (B.var[2 * i] != B.var[2 * j]) return B.var[2 * i] < B.var[2 * j];
    return B.var[2 * i + 1] < B.var[2 * j + 1];
}

template <typename T>
std::vector<T> permute(const std::vector<T>& v, const std::vector<uint32_t>& order, size_t stride) {
    std::vector<T> out(v.size());
    for (size_t k = 0; k < order.size(); ++k) {
        std::copy_n(v.begin() + order[k] * stride, stride, out.begin() + k * stride);
    }
    return out;
}

}  // namespace

extern "C" {

ckf_status ckf_read_measurements_csv(const char* path, int do_sort, ckf_measurement_buffer** out) {
    return guarded([&] {
        if (path == nullptr || out == nullptr) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null argument");
        const std::string p(path);
        std::ifstream in(p, std::ios::binary);
        if (!in) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, p + ": cannot open");
        std::string line;
        if (!std::getline(in, line)) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, p + ": empty file");
        if (!line.empty() && line.back() == '\r') line.pop_back();
        const std::vector<std::string> head = split(line);
        int where[kNCols];
        for (int c = 0; c < kNCols; ++c) {
            const auto it = std::find(head.begin(), head.end(), kColNames[c]);
            if (it == head.end()) bad(p, 1, std::string("missing column ") + kColNames[c]);
            where[c] = static_cast<int>(it - head.begin());
        }
        auto* B = new ckf_measurement_buffer();
        try {
            size_t ln = 1;
            while (std::getline(in, line)) {
                ++ln;
                if (!line.empty() && line.back() == '\r') line.pop_back();
                if (line.empty()) continue;
                const std::vector<std::string> f = split(line);
                if (f.size() != head.size()) bad(p, ln, "wrong number of fields");
                // local_key: one raw byte (the reference reads it as a char)
                const std::string& key = f[where[kLocalKey]];
                const uint8_t lk = key.empty() ? 0u : static_cast<uint8_t>(key[0]);
                float pos[2] = {0.f, 0.f}, var[2] = {0.f, 0.f};
                uint8_t sub[2] = {0u, 0u};
                uint32_t dims = 0u;
                if (lk & (1u << 1)) {
                    pos[0] = to_float(f[where[kLocal0]], p, ln);
                    var[0] = to_float(f[where[kVar0]], p, ln);
                    sub[dims++] = 0u;
                }
                if (lk & (1u << 2)) {
                    pos[1] = to_float(f[where[kLocal1]], p, ln);
                    var[1] = to_float(f[where[kVar1]], p, ln);
                    sub[dims++] = 1u;
                }
                if (dims == 0u) bad(p, ln, "local_key measures neither loc0 nor loc1");
                B->pos.insert(B->pos.end(), pos, pos + 2);
                B->var.insert(B->var.end(), var, var + 2);
                B->subspace.insert(B->subspace.end(), sub, sub + 2);
                B->dims.push_back(dims);
                B->time.push_back(to_float(f[where[kTime]], p, ln));
                B->surface.push_back(to_u64(f[where[kGeoId]], p, ln));
                B->ident.push_back(static_cast<uint32_t>(to_u64(f[where[kMeasId]], p, ln)));
            }
            const uint32_t n = static_cast<uint32_t>(B->dims.size());
            std::vector<uint32_t> order(n);
            std::iota(order.begin(), order.end(), 0u);
            if (do_sort) {
                std::stable_sort(order.begin(), order.end(),
                                 [&](uint32_t i, uint32_t j) { return meas_less(*B, i, j); });
                B->pos = permute(B->pos, order, 2);
                B->var = permute(B->var, order, 2);
                B->subspace = permute(B->subspace, order, 2);
                B->dims = permute(B->dims, order, 1);
                B->time = permute(B->time, order, 1);
                B->surface = permute(B->surface, order, 1);
                B->ident = permute(B->ident, order, 1);
            }
            B->new_index.assign(n, 0u);
            for (uint32_t k = 0; k < n; ++k) B->new_index[order[k]] = k;
        } catch (...) {
            delete B;
            throw;
        }
        *out = B;
    });
}

ckf_status ckf_measurement_buffer_view(const ckf_measurement_buffer* b, ckf_measurements* view,
                                       const uint32_t** new_index) {
    return guarded([&] {
        if (b == nullptr || view == nullptr) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null argument");
        std::memset(view, 0, sizeof(*view));
        view->size = static_cast<uint32_t>(b->dims.size());
        view->local_position = b->pos.data();
        view->local_variance = b->var.data();
        view->dimensions = b->dims.data();
        view->time = b->time.data();
        view->identifier = b->ident.data();
        view->surface_link = b->surface.data();
        view->subspace = b->subspace.data();
        if (new_index) *new_index = b->new_index.data();
    });
}

ckf_status ckf_measurement_buffer_destroy(ckf_measurement_buffer* b) {
    return guarded([&] { delete b; });
}

ckf_status ckf_write_measurements_csv(const char* path, const ckf_measurements* m) {
    return guarded([&] {
        if (path == nullptr || m == nullptr) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null argument");
        if (m->size > 0u && (m->local_position == nullptr || m->local_variance == nullptr ||
                             m->dimensions == nullptr || m->surface_link == nullptr ||
                             m->subspace == nullptr)) {
            throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "missing measurement column");
        }
        FILE* f = std::fopen(path, "wb");
        if (f == nullptr) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, std::string(path) + ": cannot open");
        std::fputs("measurement_id,geometry_id,local_key,local0,local1,phi,theta,time,"
                   "var_local0,var_local1,var_phi,var_theta,var_time\n", f);
        for (uint32_t i = 0; i < m->size; ++i) {
            uint8_t key = 0u;
            for (uint32_t d = 0; d < m->dimensions[i]; ++d) key |= static_cast<uint8_t>(1u << (m->subspace[2 * i + d] + 1u));
            const uint32_t id = m->identifier ? m->identifier[i] : i;
            const float t = m->time ? m->time[i] : 0.f;
            std::fprintf(f, "%u,%llu,%c,%.9g,%.9g,0,0,%.9g,%.9g,%.9g,0,0,0\n", id,
                         static_cast<unsigned long long>(m->surface_link[i]), static_cast<char>(key),
                         m->local_position[2 * i], m->local_position[2 * i + 1], t,
                         m->local_variance[2 * i], m->local_variance[2 * i + 1]);
        }
        if (std::fclose(f) != 0) throw ckf_error(CKF_ERR_INTERNAL, std::string(path) + ": write failed");
    });
}

}  // extern "C"
