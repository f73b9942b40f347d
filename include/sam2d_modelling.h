/*
 * sam2d_modelling.h -- the reference's own boundary, restored on top of the C ABI.
 *
 * The structs below keep the member order and types of src/src_cpu/modelling.h:13-61
 * (mod_t, wave_t, crd2d_t, geo2d_t, job_t), so code written against the
 * reference's globals links against this host layer unchanged.  The entry
 * points mirror the reference's free functions (same names, same meaning);
 * where the reference calls MPI::COMM_WORLD.Abort(code) these return the code
 * (a negative status) instead, and the driver exits with it.
 */
#ifndef SAM2D_MODELLING_H
#define SAM2D_MODELLING_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ModelInfo {
    int nx, nz;
    float dx, dz;
    char velfile[224];
    char densfile[224];
    float **vel2d, **rho2d;   /* [nx][nz] */
    float vmax, vmin;
} mod_t;

typedef struct Wavelet {
    float rec_len, fd_dt_sec, real_dt_sec;
    float dom_freq;
} wave_t;

typedef struct Coord2d {
    float x, z;
} crd2d_t;

typedef struct Geo2d {
    int nsrc, *nrec;
    crd2d_t *src2d_sp;
    crd2d_t **rec2d_sp;
} geo2d_t;

typedef struct Job {
    int adj, geomflag, fdop, frsf;
    char jbtype[8], shotfile[224], geomfile[224], logfile[224];
    char jobname[256], tmppath[512];
} job_t;

/* Process globals, as src/src_cpu/main.cpp:47-50 defines them. */
extern mod_t *mod_sp;
extern wave_t *wave_sp;
extern geo2d_t *geo2d_sp;
extern job_t *job_sp;

/* Options of the drop-in that the reference fixes at compile time or lacks. */
typedef struct sam2d_run_options {
    int npml;             /* sponge width (reference: 32, cpumodelling.cpp:15)      */
    int per_shot_files;   /* 1: <job>_shot<k>_seismogram.bin per shot; 0: the
                             reference's single <job>_seismogram.bin, rewritten by
                             every shot (cpumodelling.cpp:370-374)                  */
    int write_restart;    /* 1: write <job>checkpoint_rank0.txt after each shot, the
                             file a "Restart" job reads (cpumodelling.cpp:295-316)  */
} sam2d_run_options;
extern sam2d_run_options sam2d_opts;

/* json_parser.cpp:13-37 */
void set_job_name(const char *job);
/* json_parser.cpp:40-83 + 122-300: parse the card into the globals. */
int read_job_card(const char *job);
/* create_geometry.cpp:15-146 (shot file, geomflag = 1) */
int read_source2d(const char *shot_file);
/* read_receiver.cpp:14-91 (explicit geometry, geomflag = 0) */
int read_recv2d(const char *recv_file);
/* model.cpp:12-45 */
int read2d(const char *file, float **arr, int nx, int nz);
/* model.cpp:47-64 */
void minmax2d(float **arr, int nx, int nz, float *vmin, float *vmax);
/* create_geometry.cpp:446-459: shot 0 as "z x" lines, source first. */
int write_geometry_dump(const char *path);
/* check_stability.cpp:8-72 */
int check_stability(void);
/* modelling_module (modelling.h:106, cpumodelling.cpp:158-383) on the GPU. */
int modelling_module(int my_nsrc, geo2d_t *mygeo2d_sp);

#ifdef __cplusplus
}
#endif
#endif /* SAM2D_MODELLING_H */
