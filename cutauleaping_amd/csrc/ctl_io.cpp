// ctl_io.cpp -- dynamics writers of libctl.
//
// WriteDynamics2 (2phase_n-tau-leaping.cu:3431-3469): one text file
// <prefix>_<gid> per simulation (gid = local index, T.cu:3450-3455); row i is
// the sample time in default ostream formatting, then a TAB and the count of
// each sampled species, then '\n'.  PrintDynamics2 (T.cu:3472-3507) prints
// the same rows on stdout with an empty line after every simulation.
#include <cstdio>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

#include "ctl_internal.h"

extern "C" int ctl_write_dynamics(const ctl_sim* s, const char* path_prefix, int to_stdout) {
    ctl::clear_error();
    if (!s || (!path_prefix && !to_stdout)) return ctl::fail(CTL_E_ARGS, "null argument");
    int n, K, SS;
    const double* ts;
    int rc = ctl::sim_dims(s, &n, &K, &SS, &ts);
    if (rc) return rc;
    std::vector<uint32_t> buf((size_t)n * K * SS);
    rc = ctl_sim_samples(s, buf.data(), buf.size(), 1);
    if (rc) return rc;
    // The time column is identical for every simulation: format it once the
    // way `ostream << double` does (%g, 6 significant digits).
    std::vector<std::string> tcol(K);
    for (int i = 0; i < K; ++i) {
        std::ostringstream o;
        o << ts[i];
        tcol[i] = o.str();
    }
    std::string text;
    char num[16];
    for (int g = 0; g < n; ++g) {
        text.clear();
        const uint32_t* row = buf.data() + (size_t)g * K * SS;
        for (int i = 0; i < K; ++i) {
            text += tcol[i];
            for (int sk = 0; sk < SS; ++sk) {
                int len = std::snprintf(num, sizeof num, "\t%u", row[(size_t)i * SS + sk]);
                text.append(num, (size_t)len);
            }
            text += '\n';
        }
        if (to_stdout) {
            text += '\n';
            std::fwrite(text.data(), 1, text.size(), stdout);
        } else {
            std::string name = std::string(path_prefix) + "_" + std::to_string(g);
            FILE* f = std::fopen(name.c_str(), "wb");
            if (!f) return ctl::fail(CTL_E_ARGS, "cannot write " + name);
            std::fwrite(text.data(), 1, text.size(), f);
            std::fclose(f);
        }
    }
    if (to_stdout) std::fflush(stdout);
    return 0;
}
