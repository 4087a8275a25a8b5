// tm_io.cpp -- the output layer: the reference's FRAME / PRELIM / sweep CSV
// text, binary field snapshots with a CSV converter, and the burn-in CSV reader.
//
// Text is formatted exactly as the reference's std::ostream does it: a double
// with precision p in the default float field prints as printf("%.*g", p, x)
// (precision 6 unless the reference sets another), integers in decimal.  Large
// frames are formatted by several host threads, each on a contiguous block of
// cells, and written in order.
#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "tm/tm.h"
#include "tm_io.h"

namespace {

void put_double(std::string& s, double x, int digits)
{
    char buf[40];
    const int n = std::snprintf(buf, sizeof buf, "%.*g", digits, x);
    s.append(buf, (size_t)n);
}

void put_int(std::string& s, int64_t v)
{
    char buf[24];
    const int n = std::snprintf(buf, sizeof buf, "%lld", (long long)v);
    s.append(buf, (size_t)n);
}

int worker_count(int64_t rows)
{
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int64_t by_size = std::max<int64_t>(1, rows / 65536);
    return (int)std::min<int64_t>({(int64_t)hw, (int64_t)16, by_size});
}

// Frame rows "a,i,f0,...,f(nf-1)\n" for cells [i0, i1) of one frame; get(f, i)
// returns field f at cell i.
template <typename Get>
void frame_rows(std::string& out, const Get& get, int nf, int64_t i0, int64_t i1, double a, int digits)
{
    std::string astr;
    put_double(astr, a, digits);
    out.reserve((size_t)(i1 - i0) * (size_t)(12 + 12 * nf));
    for (int64_t i = i0; i < i1; ++i) {
        out += astr;
        out += ',';
        put_int(out, i);
        for (int f = 0; f < nf; ++f) {
            out += ',';
            put_double(out, get(f, i), digits);
        }
        out += '\n';
    }
}

// save_frame's layout (rietkerk_bjork_FKE.cpp:2849-2862): header, then one row per cell.
template <typename Get>
int write_frame(const char* path, const Get& get, int nf, int64_t N, double a, const char* header, int digits)
{
    FILE* fp = std::fopen(path, "wb");
    if (!fp) return -1;
    if (header) std::fputs(header, fp);
    const int nw = worker_count(N);
    std::vector<std::string> parts((size_t)nw);
    std::vector<std::thread> th;
    for (int w = 0; w < nw; ++w)
        th.emplace_back([&, w]() {
            const int64_t i0 = N * w / nw, i1 = N * (w + 1) / nw;
            frame_rows(parts[(size_t)w], get, nf, i0, i1, a, digits);
        });
    for (auto& t : th) t.join();
    bool ok = true;
    for (const auto& p : parts) ok = ok && std::fwrite(p.data(), 1, p.size(), fp) == p.size();
    ok = (std::fclose(fp) == 0) && ok;
    return ok ? 0 : -1;
}

struct SnapHeader {           // 64 bytes, little-endian, then nf*L*L values
    char magic[8];            // "TMSNAP1"
    int32_t model, precision, nf, L;
    uint32_t seed, instance_id;
    int64_t step;
    int32_t params_index, reserved0;
    uint64_t data_bytes;
    uint64_t reserved1;
};
static_assert(sizeof(SnapHeader) == 64, "snapshot header is 64 bytes");
const char kSnapMagic[8] = {'T', 'M', 'S', 'N', 'A', 'P', '1', '\0'};

}  // namespace

namespace tmio {

int write_snapshot(const char* path, const SnapInfo& info, const void* data, size_t bytes)
{
    SnapHeader h{};
    std::memcpy(h.magic, kSnapMagic, 8);
    h.model = info.model;
    h.precision = info.precision;
    h.nf = info.nf;
    h.L = info.L;
    h.seed = info.seed;
    h.instance_id = info.instance_id;
    h.step = info.step;
    h.params_index = info.params_index;
    h.data_bytes = bytes;
    FILE* fp = std::fopen(path, "wb");
    if (!fp) return -1;
    bool ok = std::fwrite(&h, sizeof h, 1, fp) == 1 && std::fwrite(data, 1, bytes, fp) == bytes;
    ok = (std::fclose(fp) == 0) && ok;
    return ok ? 0 : -1;
}

int read_snapshot(const char* path, SnapInfo* info, std::vector<char>* data)
{
    FILE* fp = std::fopen(path, "rb");
    if (!fp) return -1;
    SnapHeader h;
    bool ok = std::fread(&h, sizeof h, 1, fp) == 1 && std::memcmp(h.magic, kSnapMagic, 8) == 0 &&
              (h.precision == TM_F64 || h.precision == TM_F32) && h.nf >= 1 && h.nf <= 3 && h.L > 0;
    if (ok) {
        const uint64_t want = (uint64_t)h.nf * (uint64_t)h.L * (uint64_t)h.L * (h.precision == TM_F64 ? 8u : 4u);
        ok = h.data_bytes == want;
        if (ok) {
            data->resize((size_t)want);
            ok = std::fread(data->data(), 1, (size_t)want, fp) == (size_t)want;
        }
    }
    std::fclose(fp);
    if (!ok) return -1;
    *info = SnapInfo{h.model, h.precision, h.nf, h.L, h.seed, h.instance_id, h.step, h.params_index};
    return 0;
}

}  // namespace tmio

extern "C" {

int tm_format_frame_csv(const char* path, const double* fields, int nf, int L, double a, const char* header,
                        int digits)
{
    if (!path || !fields || nf < 1 || L < 1 || digits < 1 || digits > 17) return -1;
    const int64_t N = (int64_t)L * L;
    return write_frame(path, [&](int f, int64_t i) { return fields[f * N + i]; }, nf, N, a, header, digits);
}

int tm_write_prelim_csv(const char* path, const char* header, double a, int r, int g, const double* t,
                        const double* stats, int n_rows, int n_cols)
{
    if (!path || (n_rows > 0 && (!t || !stats)) || n_rows < 0 || n_cols < 0) return -1;
    std::string s;
    if (header) s += header;
    for (int i = 0; i < n_rows; ++i) {
        // save_prelimframe (rietkerk_bjork_FKE.cpp:2754-2762): a, r+1, g, t, then every
        // statistic followed by a comma
        put_double(s, a, 6);
        s += ',';
        put_int(s, (int64_t)r + 1);
        s += ',';
        put_int(s, g);
        s += ',';
        put_double(s, t[i], 6);
        s += ',';
        for (int c = 0; c < n_cols; ++c) {
            put_double(s, stats[(int64_t)i * n_cols + c], 6);
            s += ',';
        }
        s += '\n';
    }
    FILE* fp = std::fopen(path, "wb");
    if (!fp) return -1;
    const bool ok = std::fwrite(s.data(), 1, s.size(), fp) == s.size();
    return (std::fclose(fp) == 0 && ok) ? 0 : -1;
}

int tm_write_sweep_csv(const char* path, const char* header, const double* rows, int n_rows, int n_cols)
{
    if (!path || n_rows < 0 || n_cols < 1 || (n_rows > 0 && !rows)) return -1;
    // first_order_critical_exp_delta_stochastic_MultiSp (rietkerk_bjork_FKE.cpp:4864-4873):
    // the header line, then the rows with precision 16, comma-separated
    std::string s;
    if (header) {
        s += header;
        s += '\n';
    }
    for (int i = 0; i < n_rows; ++i) {
        for (int c = 0; c < n_cols; ++c) {
            if (c) s += ',';
            put_double(s, rows[(int64_t)i * n_cols + c], 16);
        }
        s += '\n';
    }
    FILE* fp = std::fopen(path, "wb");
    if (!fp) return -1;
    const bool ok = std::fwrite(s.data(), 1, s.size(), fp) == s.size();
    return (std::fclose(fp) == 0 && ok) ? 0 : -1;
}

int tm_snapshot_to_frame_csv(const char* snap, const char* csv, double a, const char* header, int digits)
{
    if (!snap || !csv || digits < 1 || digits > 17) return -1;
    tmio::SnapInfo h;
    std::vector<char> data;
    if (tmio::read_snapshot(snap, &h, &data) != 0) return -1;
    const int64_t N = (int64_t)h.L * h.L;
    if (h.precision == TM_F64) {
        const double* v = (const double*)data.data();
        return write_frame(csv, [&](int f, int64_t i) { return v[f * N + i]; }, h.nf, N, a, header, digits);
    }
    const float* v = (const float*)data.data();
    return write_frame(csv, [&](int f, int64_t i) { return (double)v[f * N + i]; }, h.nf, N, a, header, digits);
}

int tm_read_frame_csv(const char* path, int L, int nf, const char* const* cols, int n_cols, double* out)
{
    if (!path || !cols || !out || L < 1 || nf < 1 || n_cols < 1) return -1;
    FILE* fp = std::fopen(path, "rb");
    if (!fp) return -1;
    std::string text;
    {
        char buf[1 << 16];
        size_t n;
        while ((n = std::fread(buf, 1, sizeof buf, fp)) > 0) text.append(buf, n);
        std::fclose(fp);
    }
    // Header: comma-separated names with all white space removed
    // (rietkerk_bjork_FKE.cpp:700-712).
    size_t pos = text.find('\n');
    const std::string head = text.substr(0, pos == std::string::npos ? text.size() : pos);
    std::vector<std::string> names;
    {
        std::string cur;
        for (char ch : head) {
            if (ch == ',') {
                names.push_back(cur);
                cur.clear();
            } else if (!std::isspace((unsigned char)ch)) {
                cur += ch;
            }
        }
        names.push_back(cur);
    }
    // Destination field of each requested column: W(x;t) and O(x;t) go to the
    // last two fields, any other column to (its CSV position - 2)
    // (rietkerk_bjork_FKE.cpp:738-745).
    std::vector<int> dest(names.size(), -1);
    for (int k = 0; k < n_cols; ++k) {
        int idx = -1;
        for (size_t j = 0; j < names.size(); ++j)
            if (names[j] == cols[k]) idx = (int)j;   // the last match, as a map keyed by name keeps it
        if (idx < 0) return -2;                      // a requested column is missing
        const std::string c = cols[k];
        const int f = c == "W(x;t)" ? nf - 2 : c == "O(x;t)" ? nf - 1 : idx - 2;
        if (f < 0 || f >= nf) return -1;
        dest[(size_t)idx] = f;
    }
    const int64_t N = (int64_t)L * L;
    int64_t row = 0;
    while (pos != std::string::npos && row < N) {
        const size_t start = pos + 1;
        if (start >= text.size()) break;
        size_t end = text.find('\n', start);
        const size_t stop = end == std::string::npos ? text.size() : end;
        size_t c0 = start;
        for (int col = 0; c0 <= stop; ++col) {
            size_t c1 = text.find(',', c0);
            if (c1 == std::string::npos || c1 > stop) c1 = stop;
            if (col < (int)dest.size() && dest[(size_t)col] >= 0) {
                // std::stod on the cell; a cell that does not convert reads as 0
                const std::string cell = text.substr(c0, c1 - c0);
                char* e = nullptr;
                double v = std::strtod(cell.c_str(), &e);
                if (e == cell.c_str()) v = 0.0;
                out[dest[(size_t)col] * N + row] = v;
            }
            if (c1 == stop) break;
            c0 = c1 + 1;
        }
        ++row;
        pos = end;
    }
    return (int)std::min<int64_t>(row, N);
}

}  // extern "C"
