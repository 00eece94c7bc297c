// Reference-side harness (solver).  Built only where the reference sources
// exist, by oracle/Makefile.ref, into oracle/_ref/libdsmcref.so.  It drives
// the reference's own System (system.cpp:22-47, 241-382), MoleculeMover
// (moleculemover.cpp:50-126) and Cell (cell.cpp:42-111) through a C entry
// surface so that tests/golden/make_golden.py can record golden vectors, and
// wraps the configured surface collider (colliderbase.h:13) to log every wall
// hit.  It is a checker, never product code.
#include <mpi.h>
#include <system.h>
#include <grid.h>
#include <cell.h>
#include <moleculemover.h>
#include <colliderbase.h>
#include <random.h>
#include <settings.h>
#include <cvector.h>
#include <statisticssampler.h>
#include <dsmc_io.h>
#include <unistd.h>
#include <climits>
#include <cstring>
#include <vector>

namespace {

struct RecordingCollider : public ColliderBase {
    ColliderBase *inner = nullptr;
    System *sys = nullptr;
    std::vector<long> voxel, particle;
    std::vector<double> vin, vout;
    void collide(Random *rnd, data_type *v, float *n, float *t1, float *t2, bool pd) override {
        Grid *g = sys->world_grid;
        long hidx = (long)(n - g->normals) / 3;
        long hy = g->ny, hz = g->nz;
        long i = hidx / (hy * hz), j = (hidx / hz) % hy, k = hidx % hz;
        long gx = g->global_nx, gy = g->global_ny, gz = g->global_nz;
        voxel.push_back((i % gx) * gy * gz + (j % gy) * gz + (k % gz));
        particle.push_back((long)(v - sys->v.data()) / 3);
        for (int a = 0; a < 3; a++) vin.push_back(v[a]);
        inner->collide(rnd, v, n, t1, t2, pd);
        for (int a = 0; a < 3; a++) vout.push_back(v[a]);
        num_collisions = inner->num_collisions;
    }
};

struct Handle {
    Settings *settings;
    System *sys;
    RecordingCollider *rec;
    StatisticsSampler *sampler = nullptr;
};

void ensure_mpi() {
    int inited = 0;
    MPI_Initialized(&inited);
    if (!inited) MPI_Init(nullptr, nullptr);
}

}  // namespace

extern "C" {

void *ref_sys_create(const char *workdir, const char *ini) {
    ensure_mpi();
    char old[4096];
    if (!getcwd(old, sizeof old)) return nullptr;
    if (chdir(workdir) != 0) return nullptr;
    Handle *h = new Handle;
    h->settings = new Settings(ini);
    h->sys = new System();
    h->sys->initialize(h->settings, 0);
    h->sys->t0 = 0;
    h->rec = new RecordingCollider;
    h->rec->inner = h->sys->mover->surface_collider;
    h->rec->sys = h->sys;
    h->rec->num_collisions = 0;
    h->sys->mover->surface_collider = h->rec;
    if (chdir(old) != 0) return nullptr;
    return h;
}

long ref_sys_num(void *hp) { return ((Handle *)hp)->sys->num_molecules_local; }

void ref_sys_get_state(void *hp, double *rv6) {
    System *s = ((Handle *)hp)->sys;
    for (long n = 0; n < s->num_molecules_local; n++)
        for (int a = 0; a < 3; a++) { rv6[6 * n + a] = s->r[3 * n + a]; rv6[6 * n + 3 + a] = s->v[3 * n + a]; }
}

void ref_sys_set_state(void *hp, const double *rv6) {
    System *s = ((Handle *)hp)->sys;
    for (long n = 0; n < s->num_molecules_local; n++)
        for (int a = 0; a < 3; a++) { s->r[3 * n + a] = rv6[6 * n + a]; s->v[3 * n + a] = rv6[6 * n + 3 + a]; }
}

// The body of System::move() (system.cpp:107-120) with a caller-chosen dt.
void ref_sys_move_all(void *hp, double dt) {
    System *s = ((Handle *)hp)->sys;
    CVector system_length(s->length[0], s->length[1], s->length[2]);
    for (int n = 0; n < s->num_molecules_local; n++) s->mover->move_molecule(n, dt, s->rnd, 0);
    for (int n = 0; n < s->num_molecules_local; n++) s->mover->apply_periodic_boundary_conditions(n, s->r, system_length);
}

void ref_sys_step(void *hp) { ((Handle *)hp)->sys->step(); }

// The reference's StatisticsSampler (statisticssampler.cpp:13-48) on the
// statistics/ files System::initialize opened in the work directory
// (dsmc_io.cpp:33-46): sample() after each step as main.cpp:58-64 does,
// finalize() and the files' close (dsmc_io.cpp:160-182) at the end.
void ref_sys_stats_enable(void *hp) {
    Handle *h = (Handle *)hp;
    if (!h->sampler) h->sampler = new StatisticsSampler(h->sys);
}

void ref_sys_stats_sample(void *hp) { ((Handle *)hp)->sampler->sample(); }

void ref_sys_stats_finalize(void *hp) {
    Handle *h = (Handle *)hp;
    h->sampler->finalize();
    h->sys->io->finalize();
}

// DSMC_IO::save_state_to_movie_file (dsmc_io.cpp:50-80) as main.cpp:57-59
// calls it before each step: in the work directory, so the frame goes to
// movie_files/movie0000.bin (flushed for the caller to read).
int ref_sys_movie_frame(void *hp, const char *workdir) {
    Handle *h = (Handle *)hp;
    char old[4096];
    if (!getcwd(old, sizeof old) || chdir(workdir) != 0) return -1;
    h->sys->io->save_state_to_movie_file();
    if (h->sys->io->movie_file_open) h->sys->io->movie_file->flush();
    return chdir(old) == 0 ? 0 : -1;
}

// out: steps, t, collisions, wall collisions, count_periodic[3], log length
void ref_sys_counters(void *hp, double *out) {
    Handle *h = (Handle *)hp;
    System *s = h->sys;
    out[0] = s->steps;
    out[1] = s->t;
    out[2] = (double)s->collisions;
    out[3] = (double)h->rec->voxel.size();
    for (int a = 0; a < 3; a++) out[4 + a] = (double)s->mover->count_periodic[a];
    out[7] = (double)h->rec->inner->num_collisions;
}

long ref_sys_log_size(void *hp) { return (long)((Handle *)hp)->rec->voxel.size(); }

void ref_sys_log(void *hp, long *voxel, long *particle, double *vin, double *vout) {
    RecordingCollider *r = ((Handle *)hp)->rec;
    std::memcpy(voxel, r->voxel.data(), r->voxel.size() * sizeof(long));
    std::memcpy(particle, r->particle.data(), r->particle.size() * sizeof(long));
    std::memcpy(vin, r->vin.data(), r->vin.size() * sizeof(double));
    std::memcpy(vout, r->vout.data(), r->vout.size() * sizeof(double));
}

void ref_sys_log_clear(void *hp) {
    RecordingCollider *r = ((Handle *)hp)->rec;
    r->voxel.clear(); r->particle.clear(); r->vin.clear(); r->vout.clear();
}

int ref_sys_num_active(void *hp) { return (int)((Handle *)hp)->sys->active_cells.size(); }

// Per active cell (in the reference's active order): global index, count,
// vr_max, collision_rest; members receives the concatenated membership lists.
void ref_sys_cells(void *hp, int *index, int *count, double *vr_max, double *rest, int *members) {
    System *s = ((Handle *)hp)->sys;
    long off = 0;
    for (size_t c = 0; c < s->active_cells.size(); c++) {
        Cell *cell = s->active_cells[c];
        index[c] = cell->index;
        count[c] = cell->num_molecules;
        vr_max[c] = cell->vr_max;
        rest[c] = cell->collision_rest;
        if (members)
            for (int m = 0; m < cell->num_molecules; m++) members[off++] = cell->molecules[m];
    }
}

void ref_sys_cell_volumes(void *hp, int *index, double *volume, double *coeff) {
    System *s = ((Handle *)hp)->sys;
    for (size_t c = 0; c < s->active_cells.size(); c++) {
        index[c] = s->active_cells[c]->index;
        volume[c] = s->active_cells[c]->volume;
        coeff[c] = s->active_cells[c]->collision_coefficient;
    }
}

}
