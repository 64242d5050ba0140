/*
 * hfx_serve.h -- continuous-batching SCF server (no reference counterpart:
 * it generalises the reference's sequential bond-length scan, run_bl_scan in
 * include/run.hpp:300-423 of KLYYCYNN/hartree_fock_cpp, which solves one point
 * at a time with a fresh ERI_GPU per point).
 *
 * Jobs are admitted as in-flight slots and HBM budget free up; each admitted
 * job runs its SCF on its own worker thread and HIP stream, so the device work
 * of different jobs overlaps.  Admission is by the job's HBM footprint: stored
 * ERIs when they fit the job's share of the budget, else direct; a job that
 * does not fit waits while smaller queued jobs may go ahead of it.
 */
#ifndef HFX_SERVE_H
#define HFX_SERVE_H

#include <stddef.h>

#include "hfx_scf.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hfx_server hfx_server;

typedef struct {
    int max_inflight;      /* concurrent jobs (worker threads); >= 1 */
    size_t hbm_budget;     /* bytes for all in-flight jobs; 0 = unlimited */
    int allow_skip_ahead;  /* 1: a smaller queued job may start while the head job waits for memory */
} hfx_server_opts;

typedef struct {
    int natom;
    const char* symbols;       /* natom NUL-separated element symbols */
    const double* xyz;         /* 3*natom bohr */
    const char* basis_set;
    const char* basis_path;
    hfx_scf_opts scf;          /* eri_mode is decided at admission unless HFX_MODE_STORED/DIRECT is forced */
    const double* P0;          /* optional initial density (K x K), copied at submit */
} hfx_job_desc;

typedef struct {
    int status;                /* 0 ok, < 0 error (message in `error`) */
    int done;
    hfx_scf_result scf;
    int K;
    size_t device_bytes;       /* footprint charged at admission */
    double t_submit, t_admit, t_done;   /* seconds on the server's monotonic clock */
    char error[256];
} hfx_job_result;

int hfx_server_create(const hfx_server_opts* opts, hfx_server** out);
/* Stops accepting, finishes every queued and running job, frees the server. */
void hfx_server_destroy(hfx_server* s);

/* Queues a job; returns its id (0, 1, 2, ... in submission order). */
int hfx_server_submit(hfx_server* s, const hfx_job_desc* job, long long* job_id);

/* Result of a job; wait != 0 blocks until it is done.  Returns 1 when the job is
 * done (result filled), 0 when still pending (wait == 0), < 0 on a bad id. */
int hfx_server_result(hfx_server* s, long long job_id, int wait, hfx_job_result* out);

/* Blocks until every submitted job is done. */
int hfx_server_drain(hfx_server* s);

/* Seconds on the server's clock (for arrival scheduling by the caller). */
double hfx_server_now(hfx_server* s);

/* Peak concurrent jobs and peak charged bytes seen so far. */
int hfx_server_stats(hfx_server* s, int* peak_inflight, size_t* peak_bytes, long long* n_done);

/* Optional density output of a finished job (K x K). */
int hfx_server_density(hfx_server* s, long long job_id, double* P);

#ifdef __cplusplus
}
#endif

#endif /* HFX_SERVE_H */
