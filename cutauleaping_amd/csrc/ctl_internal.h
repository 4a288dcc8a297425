// ctl_internal.h -- shared declarations of libctl (host side and kernels).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/ctl.h"

namespace ctl {

// Thread-local error reporting behind ctl_last_error().
int fail(int code, const std::string& msg);
void clear_error();

// Packed entry formats of the compiled model tables.
//   left / fear entry : species (low 16 bits) | coefficient << 16 (signed for fear)
//   rav entry         : reaction (low 16 bits) | coefficient << 16 (signed)
inline uint32_t pack_entry(int idx, int coef) {
    return (uint32_t)(idx & 0xffff) | ((uint32_t)(coef & 0xffff) << 16);
}
constexpr int kMaxIndex = 65535;    // S, R limit of the packed format
constexpr int kMaxCoef = 32767;     // |coefficient| limit

// Fitness inputs, read_time_series (T.cu:1056-1261).
struct Fitness {
    bool loaded = false;
    uint32_t repetitions = 0;     // ts_rep (E)
    uint32_t par_repetitions = 0; // p_rep (P)
    uint32_t target_quantities = 0; // ts_numtgt
    uint32_t experiments = 0;     // D, from tts_vector
    std::vector<int8_t> thread2experiment;   // threads*blocks
    std::vector<uint32_t> series; // [sample][experiment][repetition][quantity]
};

struct Model {
    int S = 0, R = 0;
    std::vector<int32_t> left, right, feed;       // R*S, R*S, S
    std::vector<uint32_t> x0;                     // x0_rows * S
    int64_t x0_rows = 0;
    std::vector<double> c;                        // c_rows * R
    int64_t c_rows = 0;
    std::vector<double> ts;                       // K
    std::vector<int32_t> cs;                      // SS
    int64_t row_offset = 0;
    // per-simulation feeds (read_feeds T.cu:459-567): species fs[u] is reset to
    // fv[row][u] after every accepted leap; none when fs is empty
    std::vector<int32_t> fs;
    std::vector<uint32_t> fv;                     // fv_rows * fs.size()
    int64_t fv_rows = 0;
    // compiled (read_stoichiometry T.cu:122-457)
    int nL = 0, nV = 0, nF = 0;
    std::vector<uint32_t> Lptr, Lent, Fptr, Fent, Vptr, Vent;
    std::vector<uint8_t> hor, hort;
    Fitness fit;
};

int compile_model(Model& m);   // validates and fills the compiled tables

}  // namespace ctl

struct ctl_sim;
namespace ctl {
// Simulation count, samples, sampled species and the host copy of t_vector.
int sim_dims(const ctl_sim* s, int* n, int* K, int* SS, const double** ts);
}  // namespace ctl

struct ctl_model {
    ctl::Model m;
};
