// dsmc_run: the reference's main program (main.cpp:19-103) on one GPU,
// written against the C ABI in include/dsmc.h as a reference driver would
// be.  Run in (or pointed at) a reference run directory: it reads dsmc.ini
// and the world directory the ini names ("<world>/0000.bin", one rank),
// seeds or, with load_previous_state, loads state_files/state0000.bin and
// the Tocontinue counters, runs `timesteps` steps with the movie and the
// statistics sampler as configured, and writes state_files/state0000.bin,
// statistics/*.txt and Tocontinue as the reference does.
//
//   dsmc_run [run directory]
#include <sys/stat.h>
#include <unistd.h>

#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>

#include "../../include/dsmc.h"

namespace {

int die(dsmc_handle *h, const char *what, dsmc_status st) {
    std::fprintf(stderr, "dsmc_run: %s failed (status %d)%s%s\n", what, (int)st, h ? ": " : "",
                 h ? dsmc_last_error(h) : "");
    if (h) dsmc_destroy(h);
    return 1;
}

double seconds_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

int main(int argc, char **argv) {
    if (argc > 2 || (argc == 2 && (!std::strcmp(argv[1], "-h") || !std::strcmp(argv[1], "--help")))) {
        std::fprintf(stderr, "usage: dsmc_run [run directory holding dsmc.ini]\n");
        return 2;
    }
    if (argc == 2 && chdir(argv[1]) != 0) {
        std::perror("dsmc_run: run directory");
        return 2;
    }
    const auto t_start = std::chrono::steady_clock::now();
    dsmc_settings s;
    char world[1024];
    dsmc_status st = dsmc_settings_from_ini("dsmc.ini", &s, world, sizeof world);
    if (st != DSMC_OK) return die(nullptr, "reading dsmc.ini", st);
    dsmc_run_settings r;
    st = dsmc_run_settings_from_ini("dsmc.ini", &r);
    if (st != DSMC_OK) return die(nullptr, "reading the run keys of dsmc.ini", st);
    dsmc_world w;
    const std::string world_file = std::string(world) + "/0000.bin";
    st = dsmc_world_read(world_file.c_str(), &w);
    if (st != DSMC_OK) return die(nullptr, world_file.c_str(), st);
    dsmc_handle *h = nullptr;
    st = dsmc_create(&s, &w, &h);
    dsmc_world_free(&w);
    if (st != DSMC_OK) return die(h, "dsmc_create", st);
    for (const char *d : {"state_files", "movie_files", "statistics"}) (void)mkdir(d, 0755);

    // the movie is set before the particles exist, so that its rows follow
    // the seeded or loaded order (the reference's storage order)
    if (r.create_movie && r.movie_every_n_frame > 0) {
        st = dsmc_set_movie(h, "movie_files/movie0000.bin", r.movie_every_n_frame, r.movie_molecules);
        if (st != DSMC_OK) return die(h, "dsmc_set_movie", st);
    }
    // System::setup_molecules (system.cpp:398-430) or the saved state and the
    // counters of Tocontinue (main.cpp:37-55)
    double t = 0;
    unsigned long steps = 0, collisions = 0, wall_collisions = 0;
    if (r.load_previous_state) {
        st = dsmc_load_state_file(h, "state_files/state0000.bin");
        if (st != DSMC_OK) return die(h, "loading state_files/state0000.bin", st);
        std::ifstream to_continue("Tocontinue");
        to_continue >> t >> steps >> collisions >> wall_collisions;
    } else {
        int64_t n = 0;
        st = dsmc_initialize(h, &n);
        if (st != DSMC_OK) return die(h, "dsmc_initialize", st);
    }
    st = dsmc_set_counters(h, (int64_t)steps, t, collisions, wall_collisions);
    if (st != DSMC_OK) return die(h, "dsmc_set_counters", st);
    // StatisticsSampler: its properties sample every statistics_interval
    // steps (statisticalproperty.cpp:22); the run's time origin is t
    if (r.statistics_interval > 0) {
        st = dsmc_set_statistics(h, r.statistics_interval, r.sampling_bins,
                                 r.timesteps / r.statistics_interval + 2);
        if (st != DSMC_OK) return die(h, "dsmc_set_statistics", st);
    }

    // main.cpp:57-64: the movie frames and the samples happen inside dsmc_step
    const auto t_loop = std::chrono::steady_clock::now();
    st = dsmc_step(h, r.timesteps);
    if (st == DSMC_OK) st = dsmc_synchronize(h);
    if (st != DSMC_OK) return die(h, "dsmc_step", st);
    const double loop_s = seconds_since(t_loop);

    // main.cpp:66-72 and 94-96
    st = dsmc_save_state_file(h, "state_files/state0000.bin");
    if (st != DSMC_OK) return die(h, "saving state_files/state0000.bin", st);
    if (r.statistics_interval > 0) {
        st = dsmc_write_statistics(h, "statistics");
        if (st != DSMC_OK) return die(h, "writing statistics/", st);
    }
    st = dsmc_set_movie(h, nullptr, 0, 0);      // closes the movie file
    if (st != DSMC_OK) return die(h, "closing the movie", st);
    dsmc_counters c;
    st = dsmc_get_counters(h, &c);
    if (st != DSMC_OK) return die(h, "dsmc_get_counters", st);
    const double total_s = seconds_since(t_start);
    std::printf("Program finished after %.1f seconds (%.1f s stepping).\n", total_s, loop_s);
    if (loop_s > 0 && r.timesteps > 0) {
        std::printf("%g timesteps / second.\n", r.timesteps / loop_s);
        std::printf("%g k atom-timesteps / second.\n", (double)c.num_molecules * r.timesteps / (1000.0 * loop_s));
    }
    std::printf("steps %lld, collisions %llu, wall collisions %llu, molecules %lld\n", (long long)c.steps,
                (unsigned long long)c.collisions, (unsigned long long)c.wall_collisions,
                (long long)c.num_molecules);
    std::ofstream to_continue_write("Tocontinue");
    to_continue_write << c.t << " " << c.steps << " " << c.collisions << " " << c.wall_collisions;
    to_continue_write.close();
    dsmc_destroy(h);
    return 0;
}
