// ctl_cli.cpp -- the cuTauLeaping command line on libctl (main of kernel.cu:74-436).
//
//   cuTauLeaping input_folder threads blocks gpu offset output_folder prefix [fitness [force_ssa [verbose]]]
//
// Same positional arguments as the reference (kernel.cu:78-119, README.md:37):
//   fitness   "0" dynamics files, "1" ./pref_allfit, "2" stdout; omitted = on, stdout
//   force_ssa != "0" SSA only
//   verbose   "1" verbose, "2" dynamics on stdout
// The reference seeds its generator with time(NULL) (kernel.cu:176); the
// environment variable CTL_SEED, when set, fixes the Philox key instead.
// Deliberate differences: an existing output folder is used rather than
// falling back to ./ (T.cu:3133-3140); the pref_allfit file holds only the
// blocks/p_rep computed groups (kernel.cu:414 loops over all blocks).
#include <sys/stat.h>
#include <sys/types.h>

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/ctl.h"

static void usage() {
    std::printf("ERROR: please specify:\n");
    std::printf("                       the folder which contains the model\n");
    std::printf("                       the number of parallel threads\n");
    std::printf("                       the number of blocks\n");
    std::printf("                       the device's number\n");
    std::printf("                       the offset of data for this GPU\n");
    std::printf("                       the output folder\n");
    std::printf("                       the prefix for the output files\n");
    std::printf("                       the type of fitness (0 to disable)\n");
    std::printf("       for instance: cuTauLeaping ./input 128 4 1 0 output prefix 0\n");
}

// get_output_folder (T.cu:3114-3148): trailing '/', created if missing.
static std::string output_folder(const char* arg, bool verbose) {
    std::string f = arg;
    if (f.empty() || f.back() != '/') f += "/";
    struct stat sb;
    if (stat(f.c_str(), &sb) == 0 && S_ISDIR(sb.st_mode)) return f;
    mode_t mask = umask(0);
    int rc = mkdir(f.c_str(), S_IRWXU | S_IRGRP | S_IXGRP);
    umask(mask);
    if (rc != 0) {
        if (verbose) {
            std::printf("ERROR: unable to create results folder %s", f.c_str());
            std::perror("");
        }
        return "./";
    }
    return f;
}

static int die(int rc) {
    std::fprintf(stderr, "ERROR: %s\n", ctl_last_error());
    return rc;
}

int main(int argc, const char* argv[]) {
    if (argc < 8) {
        usage();
        return -1;
    }
    const std::string vflag = argc > 10 ? argv[10] : "0";
    const bool verbose = vflag == "1";
    const bool stdout_dyn = vflag == "2";
    const char* model_folder = argv[1];
    ctl_opts o;
    ctl_opts_default(&o);
    o.threads = std::atoi(argv[2]);
    o.blocks = std::atoi(argv[3]);
    o.gpu = std::atoi(argv[4]);
    o.offset = std::atoll(argv[5]);
    o.verbose = verbose;
    if (verbose) {
        std::printf(" * Required %d threads per block using %d blocks", o.threads, o.blocks);
        std::printf(" * Launching simulations on GPU %d \n", o.gpu);
        std::printf(" * Data offset: %lld to %lld threads\n", (long long)o.offset,
                    (long long)(o.offset + (long long)o.threads * o.blocks));
    }
    const std::string out_folder = output_folder(argv[6], verbose);
    const std::string prefix = argv[7];
    if (verbose) std::printf(" * Output path: %s%s.\n", out_folder.c_str(), prefix.c_str());
    bool just_fitness = true, text_fitness = true;
    if (argc > 8) {
        just_fitness = std::string(argv[8]) != "0";
        text_fitness = std::string(argv[8]) == "2";
    }
    if (argc > 9) {
        o.force_ssa = std::string(argv[9]) != "0";
        if (o.force_ssa) std::fprintf(stdout, " * tau-leaping disabled: cuTauLeaping will force SSA instead.\n");
    }
    const char* seed_env = std::getenv("CTL_SEED");
    o.seed = seed_env ? std::strtoull(seed_env, nullptr, 0) : (uint64_t)std::time(nullptr);
    if (verbose) std::printf(" * Seed: %llu\n", (unsigned long long)o.seed);
    if (o.threads <= 0 || o.blocks <= 0) {
        usage();
        return -1;
    }
    const int64_t n = (int64_t)o.threads * o.blocks;

    ctl_model* model = nullptr;
    int rc = ctl_model_load_folder(model_folder, o.offset, n, verbose, &model);
    if (rc) return die(rc);
    if (just_fitness) {
        rc = ctl_model_load_fitness(model, model_folder, o.threads, o.blocks, verbose);
        if (rc) return die(rc);
    }
    // Per-simulation feeds are compiled out of the reference (DISABLE_FEED,
    // T.cu:34); CTL_FEEDS=1 turns them on (read_feeds, T.cu:459-567).
    const char* feeds_env = std::getenv("CTL_FEEDS");
    if (feeds_env && std::string(feeds_env) != "0") {
        rc = ctl_model_load_feeds(model, model_folder, o.offset, n, verbose);
        if (rc) return die(rc);
    }
    // CTL_LANES sets lanes per simulation (default: the library's choice).
    const char* lanes_env = std::getenv("CTL_LANES");
    if (lanes_env) o.lanes_per_sim = std::atoi(lanes_env);
    ctl_sim* sim = nullptr;
    rc = ctl_sim_create(model, &o, &sim);
    if (rc) return die(rc);
    if (verbose) std::fprintf(stdout, " * 3P-tau-leaping starting on GPU%d\n", o.gpu);
    rc = ctl_sim_run(sim);
    if (rc) return die(rc);
    if (verbose) {
        ctl_stats st;
        ctl_sim_stats(sim, &st);
        std::printf(" * RUNNING TIME: %f seconds\n", st.kernel_ms / 1000.0);
    }
    const std::string dumpfile = out_folder + prefix;
    if (!just_fitness) {
        if (!stdout_dyn) {
            std::printf(" * Writing dynamics to file %s\n", dumpfile.c_str());
            std::fflush(stdout);
            rc = ctl_write_dynamics(sim, dumpfile.c_str(), 0);
        } else {
            std::fflush(stdout);
            rc = ctl_write_dynamics(sim, dumpfile.c_str(), 1);
        }
        if (rc) return die(rc);
    } else {
        if (verbose) std::printf(" * Calculating fitnesses.\n");
        ctl_model_info info;
        ctl_model_info_get(model, &info);
        // groups = blocks / p_rep, read back through the fitness buffer size
        std::ifstream pr((std::string(model_folder) + "/p_rep").c_str());
        int P = 1;
        if (pr) pr >> P;
        if (P <= 0) P = 1;
        const int groups = o.blocks / P;
        std::vector<double> fit((size_t)o.threads * groups);
        rc = ctl_fitness(sim, fit.data(), fit.size());
        if (rc) return die(rc);
        if (text_fitness) {
            for (int sw = 0; sw < groups; ++sw) {
                for (int part = 0; part < o.threads; ++part) std::cout << fit[(size_t)sw * o.threads + part] << "\t";
                std::cout << "\t";
            }
            std::cout.flush();
        } else {
            std::printf(" * Dumping fitness to file.\n");
            std::ofstream dump("./pref_allfit");
            dump << std::fixed << std::showpoint << std::setprecision(20);
            for (int sw = 0; sw < groups; ++sw) {
                for (int part = 0; part < o.threads; ++part) dump << fit[(size_t)sw * o.threads + part] << "\t";
                dump << "\n";
            }
        }
    }
    ctl_sim_destroy(sim);
    ctl_model_destroy(model);
    return 0;
}
