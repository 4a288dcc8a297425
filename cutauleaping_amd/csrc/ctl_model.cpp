// ctl_model.cpp -- input-folder parsers and model compile of libctl.
//
// Restates the reference's host-side input layer with its observable
// semantics:
//   first_pass           2phase_n-tau-leaping.cu:90-120
//   read_stoichiometry   2phase_n-tau-leaping.cu:122-457
//   read_quantities_offset 2phase_n-tau-leaping.cu:570-649
//   read_parameters_offset 2phase_n-tau-leaping.cu:717-782
//   read_instants        2phase_n-tau-leaping.cu:932-1053
//   read_time_series     2phase_n-tau-leaping.cu:1056-1261
// Tokens are TAB-separated (getline(.., '\t'), T.cu:78-82); integers are read
// like `stream >> int` (leading integer, 0 when none) and counts like
// `stream >> double` then cast to unsigned (T.cu:620-624).
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

#include "ctl_internal.h"

namespace ctl {

namespace {
thread_local std::string g_err;
}

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}
void clear_error() { g_err.clear(); }

namespace {

bool read_lines(const std::string& path, std::vector<std::string>& lines) {
    std::ifstream in(path.c_str(), std::ios::in | std::ios::binary);
    if (!in.is_open()) return false;
    std::string buf((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    lines.clear();
    size_t pos = 0;
    while (pos < buf.size()) {               // getline semantics: no empty last line
        size_t e = buf.find('\n', pos);
        if (e == std::string::npos) e = buf.size();
        size_t end = e;
        if (end > pos && buf[end - 1] == '\r') --end;
        lines.emplace_back(buf, pos, end - pos);
        pos = e + 1;
    }
    return true;
}

// getline(ss, token, '\t') splitting: a trailing TAB yields no empty token.
void split_tabs(const char* p, const char* end, std::vector<std::pair<const char*, const char*>>& out) {
    out.clear();
    if (p == end) return;
    const char* s = p;
    for (const char* q = p; q < end; ++q)
        if (*q == '\t') { out.emplace_back(s, q); s = q + 1; }
    if (s < end) out.emplace_back(s, end);
}

// `stringstream(token) >> int` : leading whitespace, optional sign, digits; 0 if none.
long parse_int(const char* s, const char* e) {
    std::string t(s, e);
    char* endp = nullptr;
    long v = std::strtol(t.c_str(), &endp, 10);
    if (endp == t.c_str()) return 0;
    return v;
}

// `stringstream(token) >> double`; ok=false when nothing parses.
double parse_double(const char* s, const char* e, bool* ok) {
    char tmp[64];
    size_t len = (size_t)(e - s);
    std::string big;
    const char* src;
    if (len < sizeof(tmp)) { std::memcpy(tmp, s, len); tmp[len] = 0; src = tmp; }
    else { big.assign(s, e); src = big.c_str(); }
    char* endp = nullptr;
    double v = std::strtod(src, &endp);
    *ok = endp != src;
    return *ok ? v : 0.0;
}

// (unsigned int) of a double the way the reference casts MX_0 values.
uint32_t count_from_double(double d) {
    if (!(d > 0.0)) return 0u;
    if (d >= 4294967295.0) return 4294967295u;
    return (uint32_t)d;
}

int read_matrix_int(const std::string& path, int S_expected, std::vector<int32_t>& out, int* rows,
                    int* cols, const char* what) {
    std::vector<std::string> lines;
    if (!read_lines(path, lines))
        return fail(CTL_E_ARGS, std::string("ERROR reading ") + what + " matrix: " + path + ": " +
                                    std::strerror(errno));
    if (lines.empty()) return fail(CTL_E_ARGS, std::string(what) + " matrix is empty: " + path);
    std::vector<std::pair<const char*, const char*>> tok;
    split_tabs(lines[0].data(), lines[0].data() + lines[0].size(), tok);
    int S = S_expected > 0 ? S_expected : (int)tok.size();
    if (S <= 0) return fail(CTL_E_ARGS, std::string(what) + " matrix has no columns: " + path);
    out.assign((size_t)lines.size() * S, 0);
    for (size_t r = 0; r < lines.size(); ++r) {
        split_tabs(lines[r].data(), lines[r].data() + lines[r].size(), tok);
        if ((int)tok.size() != S) {
            char b[256];
            std::snprintf(b, sizeof b, "%s matrix row %zu has %zu entries, expected %d", what, r,
                          tok.size(), S);
            return fail(CTL_E_ARGS, b);
        }
        for (int s = 0; s < S; ++s) out[r * S + s] = (int32_t)parse_int(tok[s].first, tok[s].second);
    }
    *rows = (int)lines.size();
    *cols = S;
    return 0;
}

}  // namespace

int compile_model(Model& m) {
    const int S = m.S, R = m.R;
    if (S <= 0 || R <= 0) return fail(CTL_E_ARGS, "model needs at least one species and one reaction");
    if (S > kMaxIndex || R > kMaxIndex) return fail(CTL_E_LIMIT, "model exceeds 65535 species or reactions");
    if ((int64_t)m.left.size() != (int64_t)R * S || (int64_t)m.right.size() != (int64_t)R * S)
        return fail(CTL_E_ARGS, "stoichiometry size mismatch");
    if (m.feed.empty()) m.feed.assign(S, 0);
    if ((int)m.feed.size() != S) return fail(CTL_E_ARGS, "M_feed length differs from the species count");
    if (m.ts.empty()) return fail(CTL_E_ARGS, "t_vector is empty");
    for (int sk = 0; sk < (int)m.cs.size(); ++sk)
        if (m.cs[sk] < 0 || m.cs[sk] >= S) return fail(CTL_E_ARGS, "cs_vector names a species out of range");
    m.Lptr.assign(R + 1, 0); m.Fptr.assign(R + 1, 0); m.Vptr.assign(S + 1, 0);
    m.Lent.clear(); m.Fent.clear(); m.Vent.clear();
    std::vector<int32_t> var((size_t)R * S), fear((size_t)R * S);
    for (int r = 0; r < R; ++r)
        for (int s = 0; s < S; ++s) {
            int l = m.left[(size_t)r * S + s], rt = m.right[(size_t)r * S + s];
            if (l < 0 || rt < 0) return fail(CTL_E_ARGS, "negative stoichiometric coefficient");
            if (l > kMaxCoef || rt > kMaxCoef) return fail(CTL_E_LIMIT, "stoichiometric coefficient too large");
            // T.cu:270-286: net change; feed-locked columns are zeroed in fear
            var[(size_t)r * S + s] = rt - l;
            fear[(size_t)r * S + s] = m.feed[s] ? 0 : rt - l;
        }
    for (int r = 0; r < R; ++r) {            // reaction-major flats, T.cu:219-232, 323-335
        m.Lptr[r] = (uint32_t)m.Lent.size();
        m.Fptr[r] = (uint32_t)m.Fent.size();
        for (int s = 0; s < S; ++s) {
            int l = m.left[(size_t)r * S + s];
            if (l) m.Lent.push_back(pack_entry(s, l));
            int f = fear[(size_t)r * S + s];
            if (f) m.Fent.push_back(pack_entry(s, f));
        }
    }
    m.Lptr[R] = (uint32_t)m.Lent.size();
    m.Fptr[R] = (uint32_t)m.Fent.size();
    for (int s = 0; s < S; ++s) {            // species-major var flat (rav), T.cu:308-320
        m.Vptr[s] = (uint32_t)m.Vent.size();
        for (int r = 0; r < R; ++r) {
            int v = var[(size_t)r * S + s];
            if (v) m.Vent.push_back(pack_entry(r, v));
        }
    }
    m.Vptr[S] = (uint32_t)m.Vent.size();
    m.nL = (int)m.Lent.size(); m.nV = (int)m.Vent.size(); m.nF = (int)m.Fent.size();
    // HOR / HORtype, T.cu:402-432.  Reaction order is accumulated in an
    // unsigned char; the type is set with a higher order and then raised to
    // any larger left coefficient of the species (unbraced second if).
    std::vector<uint8_t> order(R, 0);
    for (int r = 0; r < R; ++r) {
        uint8_t tot = 0;
        for (int s = 0; s < S; ++s) tot = (uint8_t)(tot + m.left[(size_t)r * S + s]);
        order[r] = tot;
    }
    m.hor.assign(S, 0); m.hort.assign(S, 0);
    for (int s = 0; s < S; ++s)
        for (int r = 0; r < R; ++r) {
            int l = m.left[(size_t)r * S + s];
            if (l > 0 && order[r] > m.hor[s]) { m.hor[s] = order[r]; m.hort[s] = (uint8_t)l; }
            if ((int)m.hort[s] < l) m.hort[s] = (uint8_t)l;
        }
    return 0;
}

}  // namespace ctl

using namespace ctl;

extern "C" {

const char* ctl_last_error(void) { return g_err.c_str(); }
const char* ctl_version(void) { return "ctl 0.1 (gfx950)"; }

void ctl_opts_default(ctl_opts* o) {
    std::memset(o, 0, sizeof(*o));
    o->threads = 64;
    o->blocks = 1;
    o->gpu = 0;
    o->offset = 0;
    o->force_ssa = 0;
    o->nc = 10;
    o->eps = 0.03f;
    o->ssa_steps = 100;
    o->ssa_factor = 10.0f;
    o->seed = 0x5EEDC0DEull;
}

int ctl_model_from_arrays(const ctl_model_desc* d, ctl_model** out) {
    clear_error();
    if (!d || !out) return fail(CTL_E_ARGS, "null argument");
    *out = nullptr;
    if (d->S <= 0 || d->R <= 0 || !d->left || !d->right || !d->x0 || !d->c || !d->t_vector || d->K <= 0 ||
        d->x0_rows <= 0 || d->c_rows <= 0)
        return fail(CTL_E_ARGS, "incomplete model description");
    ctl_model* h = new ctl_model();
    Model& m = h->m;
    m.S = d->S; m.R = d->R;
    m.left.assign(d->left, d->left + (size_t)d->R * d->S);
    m.right.assign(d->right, d->right + (size_t)d->R * d->S);
    if (d->feed) m.feed.assign(d->feed, d->feed + d->S);
    m.x0.assign(d->x0, d->x0 + (size_t)d->x0_rows * d->S);
    m.x0_rows = d->x0_rows;
    m.c.assign(d->c, d->c + (size_t)d->c_rows * d->R);
    m.c_rows = d->c_rows;
    m.ts.assign(d->t_vector, d->t_vector + d->K);
    if (d->cs_vector && d->SS > 0) m.cs.assign(d->cs_vector, d->cs_vector + d->SS);
    else for (int s = 0; s < d->S; ++s) m.cs.push_back(s);
    int rc = compile_model(m);
    if (rc) { delete h; return rc; }
    *out = h;
    return 0;
}

int ctl_model_load_folder(const char* dir, int64_t offset, int64_t n, int verbose, ctl_model** out) {
    clear_error();
    if (!dir || !out) return fail(CTL_E_ARGS, "null argument");
    *out = nullptr;
    if (offset < 0) return fail(CTL_E_ARGS, "negative offset");
    std::string folder(dir);
    ctl_model* h = new ctl_model();
    Model& m = h->m;
    auto bail = [&](int rc) { delete h; return rc; };
    int rows = 0, cols = 0, rc;
    // first_pass + left matrix (T.cu:90-120, 168-214)
    rc = read_matrix_int(folder + "/left_side", 0, m.left, &rows, &cols, "left");
    if (rc) return bail(rc);
    m.R = rows; m.S = cols;
    if (verbose) std::printf(" * Found %d reactions\n * Found %d species\n", m.R, m.S);
    int rrows = 0, rcols = 0;
    rc = read_matrix_int(folder + "/right_side", m.S, m.right, &rrows, &rcols, "right");
    if (rc) return bail(rc);
    if (rrows != m.R) return bail(fail(CTL_E_ARGS, "WARNING: left and right matrices have different size, aborting"));
    // M_feed: optional, first line; non-zero entries lock the species (T.cu:145-163)
    {
        std::vector<std::string> lines;
        m.feed.assign(m.S, 0);
        if (read_lines(folder + "/M_feed", lines) && !lines.empty()) {
            std::vector<std::pair<const char*, const char*>> tok;
            split_tabs(lines[0].data(), lines[0].data() + lines[0].size(), tok);
            for (size_t s = 0; s < tok.size() && (int)s < m.S; ++s)
                m.feed[s] = parse_int(tok[s].first, tok[s].second) != 0;
        } else if (verbose) {
            std::printf("WARNING: vector of feed species ML_feed is not specified\n");
        }
    }
    // MX_0 rows [offset, offset+n) (T.cu:570-649)
    {
        std::vector<std::string> lines;
        if (!read_lines(folder + "/MX_0", lines))
            return bail(fail(CTL_E_NOFILE, "ERROR reading quantities matrix: " + folder + "/MX_0"));
        int64_t avail = (int64_t)lines.size();
        int64_t count = n > 0 ? n : avail - offset;
        if (count <= 0 || offset + count > avail) {
            char b[200];
            std::snprintf(b, sizeof b, "ERROR: input files rows (%lld) and requested threads (%lld) do not fit",
                          (long long)avail, (long long)(offset + (n > 0 ? n : 0)));
            return bail(fail(CTL_E_ROWS, b));
        }
        m.x0.assign((size_t)count * m.S, 0);
        std::vector<std::pair<const char*, const char*>> tok;
        for (int64_t i = 0; i < count; ++i) {
            const std::string& L = lines[(size_t)(offset + i)];
            split_tabs(L.data(), L.data() + L.size(), tok);
            if ((int)tok.size() != m.S) {
                char b[200];
                std::snprintf(b, sizeof b, "ERROR! too many quantities %zu (riga %lld)", tok.size(), (long long)i);
                return bail(fail(CTL_E_WIDTH, b));
            }
            for (int s = 0; s < m.S; ++s) {
                bool ok;
                double v = parse_double(tok[s].first, tok[s].second, &ok);
                m.x0[(size_t)i * m.S + s] = count_from_double(v);
            }
        }
        m.x0_rows = count;
        m.row_offset = offset;
    }
    // c_matrix rows [offset, offset+n) (T.cu:717-782)
    {
        std::vector<std::string> lines;
        if (!read_lines(folder + "/c_matrix", lines))
            return bail(fail(CTL_E_ARGS, "ERROR reading parameters matrix: " + folder + "/c_matrix"));
        int64_t count = m.x0_rows;
        if (offset + count > (int64_t)lines.size())
            return bail(fail(CTL_E_ROWS, "ERROR: c_matrix has fewer rows than offset + simulations"));
        m.c.assign((size_t)count * m.R, 0.0);
        std::vector<std::pair<const char*, const char*>> tok;
        for (int64_t i = 0; i < count; ++i) {
            const std::string& L = lines[(size_t)(offset + i)];
            split_tabs(L.data(), L.data() + L.size(), tok);
            if ((int)tok.size() != m.R) {
                char b[200];
                std::snprintf(b, sizeof b, "ERROR! too many parameters (%zu)", tok.size());
                return bail(fail(CTL_E_ARGS, b));
            }
            for (int r = 0; r < m.R; ++r) {
                bool ok;
                m.c[(size_t)i * m.R + r] = parse_double(tok[r].first, tok[r].second, &ok);
            }
        }
        m.c_rows = count;
    }
    // t_vector (T.cu:932-991): t_max is the last entry; an unparsable line
    // repeats the previous value like a failed `stream >> double`.
    {
        std::vector<std::string> lines;
        if (!read_lines(folder + "/t_vector", lines))
            return bail(fail(CTL_E_ARGS, "ERROR reading samples vector: " + folder + "/t_vector"));
        double q = 0.0;
        for (const std::string& L : lines) {
            bool ok;
            double v = parse_double(L.data(), L.data() + L.size(), &ok);
            if (ok) q = v;
            m.ts.push_back(q);
        }
        if (m.ts.empty()) return bail(fail(CTL_E_ARGS, "t_vector is empty"));
        if (verbose) std::printf(" * %zu temporal instants required, up to t=%f.\n", m.ts.size(), m.ts.back());
    }
    // cs_vector (T.cu:997-1047): optional; default all species with a warning.
    {
        std::vector<std::string> lines;
        if (read_lines(folder + "/cs_vector", lines)) {
            for (const std::string& L : lines) {
                bool ok;
                double v = parse_double(L.data(), L.data() + L.size(), &ok);
                m.cs.push_back(ok ? (int32_t)v : (m.cs.empty() ? 0 : m.cs.back()));
            }
        } else {
            std::printf("WARNING: chemical species to sample vector not found, saving ALL species as default.\n");
            for (int s = 0; s < m.S; ++s) m.cs.push_back(s);
        }
        if (verbose) std::printf(" * %zu species will be sampled\n", m.cs.size());
    }
    rc = compile_model(m);
    if (rc) return bail(rc);
    // Identical rows collapse to one shared row (the kernels then read a
    // broadcast copy instead of a per-simulation row).
    auto collapse_u = [](std::vector<uint32_t>& v, int64_t& rows, int w) {
        for (int64_t i = 1; i < rows; ++i)
            if (std::memcmp(&v[(size_t)i * w], &v[0], sizeof(uint32_t) * w)) return;
        if (rows > 1) { v.resize(w); rows = 1; }
    };
    auto collapse_d = [](std::vector<double>& v, int64_t& rows, int w) {
        for (int64_t i = 1; i < rows; ++i)
            if (std::memcmp(&v[(size_t)i * w], &v[0], sizeof(double) * w)) return;
        if (rows > 1) { v.resize(w); rows = 1; }
    };
    collapse_u(m.x0, m.x0_rows, m.S);
    collapse_d(m.c, m.c_rows, m.R);
    *out = h;
    return 0;
}

int ctl_model_info_get(const ctl_model* h, ctl_model_info* o) {
    if (!h || !o) return fail(CTL_E_ARGS, "null argument");
    const Model& m = h->m;
    o->S = m.S; o->R = m.R; o->K = (int32_t)m.ts.size(); o->SS = (int32_t)m.cs.size();
    o->nL = m.nL; o->nV = m.nV; o->nF = m.nF;
    o->x0_rows = m.x0_rows; o->c_rows = m.c_rows; o->row_offset = m.row_offset;
    o->t_max = m.ts.back();
    return 0;
}

int ctl_model_structure(const ctl_model* h, uint8_t* hor, uint8_t* hortype, int32_t* left_flat,
                        int32_t* rav_flat, int32_t* fear_flat) {
    if (!h) return fail(CTL_E_ARGS, "null model");
    const Model& m = h->m;
    if (hor) std::memcpy(hor, m.hor.data(), m.S);
    if (hortype) std::memcpy(hortype, m.hort.data(), m.S);
    auto unpack = [](uint32_t e, int* idx, int* coef) {
        *idx = (int)(e & 0xffff);
        *coef = (int)(int16_t)(e >> 16);
    };
    if (left_flat)
        for (int r = 0; r < m.R; ++r)
            for (uint32_t h2 = m.Lptr[r]; h2 < m.Lptr[r + 1]; ++h2) {
                int s, v; unpack(m.Lent[h2], &s, &v);
                left_flat[3 * h2] = r; left_flat[3 * h2 + 1] = s; left_flat[3 * h2 + 2] = v;
            }
    if (fear_flat)
        for (int r = 0; r < m.R; ++r)
            for (uint32_t h2 = m.Fptr[r]; h2 < m.Fptr[r + 1]; ++h2) {
                int s, v; unpack(m.Fent[h2], &s, &v);
                fear_flat[3 * h2] = r; fear_flat[3 * h2 + 1] = s; fear_flat[3 * h2 + 2] = v;
            }
    if (rav_flat)
        for (int s = 0; s < m.S; ++s)
            for (uint32_t h2 = m.Vptr[s]; h2 < m.Vptr[s + 1]; ++h2) {
                int r, v; unpack(m.Vent[h2], &r, &v);
                rav_flat[3 * h2] = r; rav_flat[3 * h2 + 1] = s; rav_flat[3 * h2 + 2] = v;
            }
    return 0;
}


// Per-simulation feeds.  read_feeds (T.cu:459-567): the species whose value
// is non-zero in the LAST line of MX_feed are the fed species; rows
// [offset, offset+n) give each simulation's values for them (tab-separated,
// all S columns).  The reference compiles this out (DISABLE_FEED, T.cu:34),
// so feeds are only active when a caller loads them.  A missing MX_feed is a
// warning, not an error (T.cu:473-477): the model keeps no feeds.
int ctl_model_load_feeds(ctl_model* h, const char* dir, int64_t offset, int64_t n, int verbose) {
    clear_error();
    if (!h || !dir || offset < 0) return fail(CTL_E_ARGS, "bad arguments");
    Model& m = h->m;
    std::vector<std::string> lines;
    if (!read_lines(std::string(dir) + "/MX_feed", lines) || lines.empty()) {
        std::printf("WARNING: cannot read feed quantities\n");
        std::fflush(stdout);
        m.fs.clear(); m.fv.clear(); m.fv_rows = 0;
        return 0;
    }
    std::vector<std::pair<const char*, const char*>> tok;
    const std::string& last = lines.back();
    split_tabs(last.data(), last.data() + last.size(), tok);
    std::vector<int32_t> fs;
    for (size_t s2 = 0; s2 < tok.size(); ++s2) {
        bool ok;
        if (parse_double(tok[s2].first, tok[s2].second, &ok) != 0.0) fs.push_back((int32_t)s2);
    }
    if (fs.empty()) {
        if (verbose) std::printf(" * No species have feed\n");
        m.fs.clear(); m.fv.clear(); m.fv_rows = 0;
        return 0;
    }
    int64_t count = n > 0 ? n : (int64_t)lines.size() - offset;
    if (count <= 0 || offset + count > (int64_t)lines.size())
        return fail(CTL_E_ROWS, "MX_feed has fewer rows than offset + simulations");
    std::vector<uint32_t> fv((size_t)count * fs.size());
    for (int64_t i = 0; i < count; ++i) {
        const std::string& L = lines[(size_t)(offset + i)];
        split_tabs(L.data(), L.data() + L.size(), tok);
        if ((int)tok.size() != m.S) {
            char b[200];
            std::snprintf(b, sizeof b, "ERROR! too many quantities %zu (riga %lld)", tok.size(), (long long)i);
            return fail(CTL_E_WIDTH, b);
        }
        for (size_t u = 0; u < fs.size(); ++u) {
            bool ok;
            fv[(size_t)i * fs.size() + u] = count_from_double(parse_double(tok[fs[u]].first, tok[fs[u]].second, &ok));
        }
    }
    m.fs = fs;
    m.fv = fv;
    m.fv_rows = count;
    if (verbose) std::printf(" * %zu species have feed\n", fs.size());
    return 0;
}

int ctl_model_set_feeds(ctl_model* h, const int32_t* species, int32_t nfs, const uint32_t* values, int64_t rows) {
    clear_error();
    if (!h || nfs < 0 || (nfs > 0 && (!species || !values || rows <= 0))) return fail(CTL_E_ARGS, "bad arguments");
    Model& m = h->m;
    for (int u = 0; u < nfs; ++u) {
        if (species[u] < 0 || species[u] >= m.S) return fail(CTL_E_ARGS, "feed species out of range");
        for (int w = 0; w < u; ++w)
            if (species[w] == species[u]) return fail(CTL_E_ARGS, "feed species repeated");
    }
    m.fs.assign(species, species + nfs);
    m.fv.assign(values, values + (size_t)(nfs > 0 ? rows : 0) * nfs);
    m.fv_rows = nfs > 0 ? rows : 0;
    return 0;
}

// read_time_series (T.cu:1056-1261).  Missing files only warn in the
// reference; here p_rep is required (the fitness launch divides by it).
int ctl_model_load_fitness(ctl_model* h, const char* dir, int32_t threads, int32_t blocks, int verbose) {
    clear_error();
    if (!h || !dir) return fail(CTL_E_ARGS, "null argument");
    Model& m = h->m;
    Fitness& f = m.fit;
    f = Fitness();
    std::string folder(dir);
    std::vector<std::string> lines;
    auto first_int = [&](const char* name, uint32_t* v) -> bool {
        if (!read_lines(folder + "/" + name, lines) || lines.empty()) {
            std::fprintf(stderr, "WARNING: cannot open %s file\n", name);
            return false;
        }
        *v = (uint32_t)std::atoi(lines[0].c_str());
        return true;
    };
    first_int("ts_rep", &f.repetitions);
    if (!first_int("p_rep", &f.par_repetitions) || f.par_repetitions == 0)
        return fail(CTL_E_ARGS, "p_rep missing or zero: fitness needs the parallel repetitions");
    first_int("ts_numtgt", &f.target_quantities);
    const int64_t n = (int64_t)threads * blocks;
    f.thread2experiment.assign((size_t)n, 0);
    if (read_lines(folder + "/tts_vector", lines)) {
        int pp = 0;
        int64_t last = 0;
        for (const std::string& L : lines) {
            if (L.size() < 1) continue;
            int64_t upto = std::atoi(L.c_str());
            for (int64_t kk = last; kk < upto && kk < n; ++kk) f.thread2experiment[(size_t)kk] = (int8_t)pp;
            last = upto;
            ++pp;
        }
        f.experiments = (uint32_t)pp;
        if (last != n) {
            char b[200];
            std::snprintf(b, sizeof b, "ERROR! cannot assign threads (%lld) to initial conditions (%lld), aborting",
                          (long long)n, (long long)last);
            return fail(CTL_E_TTS, b);
        }
    } else {
        std::fprintf(stderr, "WARNING: cannot open tts_vector file\n");
    }
    const size_t cols = (size_t)f.experiments * f.repetitions * f.target_quantities;
    f.series.assign(cols * m.ts.size(), 0);
    if (read_lines(folder + "/ts_matrix", lines)) {
        size_t riga = 0;
        std::vector<std::pair<const char*, const char*>> tok;
        for (const std::string& L : lines) {
            if (L.size() < 2) continue;
            if (riga >= m.ts.size()) break;
            split_tabs(L.data(), L.data() + L.size(), tok);
            for (size_t kk = 1; kk < cols + 1 && kk < tok.size(); ++kk)
                f.series[riga * cols + kk - 1] = (uint32_t)parse_int(tok[kk].first, tok[kk].second);
            ++riga;
        }
        if (riga != m.ts.size())
            std::printf("ERROR! t_vector (%zu) and ts_matrix (%zu) have different size\n", m.ts.size(), riga);
    } else {
        std::fprintf(stderr, "WARNING: cannot open ts file (fitness unavailable)\n");
    }
    if (verbose)
        std::printf(" * Experiments: %u, repetitions: %u target quantities: %u\n", f.experiments, f.repetitions,
                    f.target_quantities);
    f.loaded = true;
    return 0;
}

void ctl_model_destroy(ctl_model* h) { delete h; }

}  // extern "C"
