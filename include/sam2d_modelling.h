/*
 * sam2d_modelling.h -- the reference's host boundary, rebuilt over the C ABI.
 *
 * The seam.  The reference swaps its modelling backend at link time behind
 * one C++ function, void modelling_module(int my_nsrc, geo2d_t *mygeo2d_sp)
 * (src/src_cpu/modelling.h:106; defined by each backend, e.g.
 * src/src_cpu/cpumodelling.cpp:158).  libsam2d_host.so defines that function
 * with the same signature and C++ linkage (declared at the end of this
 * header), so the reference's unmodified front-end objects (main,
 * json_parser, modelling_master, modelling_worker, read_receiver,
 * create_geometry, workload, model, check_stability) link against
 * libsam2d_host.so in place of cpumodelling.o and cpu_modelling_kernels.o.
 * The structs keep the member order, types and struct tags of
 * modelling.h:13-61, so both sides share the process globals mod_sp,
 * wave_sp, geo2d_sp and job_sp.
 *
 * The front end.  The other functions are this build's own versions of the
 * reference's front-end functions (same names and meaning as in
 * json_parser.cpp, create_geometry.cpp, read_receiver.cpp, model.cpp and
 * check_stability.cpp), used by SeisAcoMod2D_MI355X.exe.  They have C
 * linkage and return the reference's abort code where the reference's C++
 * functions return void and call MPI::COMM_WORLD.Abort(code), so they do not
 * replace those objects (their C names cannot collide with the reference's
 * C++ ones either).
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

/* Options of the modelling that the reference fixes at compile time or
 * lacks.  The C++ modelling_module below also takes them from the
 * environment (SAM2D_NPML, SAM2D_RESET_QUIRK, SAM2D_PER_SHOT_FILES,
 * SAM2D_WRITE_RESTART, SAM2D_SU_OUTPUT), the only channel a reference front
 * end has. */
typedef struct sam2d_run_options {
    int npml;             /* sponge width (reference: 32, cpumodelling.cpp:15)      */
    int per_shot_files;   /* 1: <job>_shot<k>_seismogram.bin per shot; 0: the
                             reference's single <job>_seismogram.bin, rewritten by
                             every shot (cpumodelling.cpp:370-374)                  */
    int write_restart;    /* 1: write <job>checkpoint_rank0.txt after each shot, the
                             file a "Restart" job reads (cpumodelling.cpp:295-316)  */
    int reset_quirk;      /* 1: each shot starts from the previous shot's final
                             wavefield with only the first nnz cells cleared, as
                             cpumodelling.cpp:343-348 does; 0: fresh state per shot */
    int su_output;        /* 1: also write the traces as an SU file (240-byte
                             SEG-Y trace headers, native-endian float samples):
                             <job>_seismogram.su, one file for the job with every
                             shot's traces (per shot with per_shot_files)           */
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

/* The output of shot `is` of geo (cpumodelling.cpp:367-374): seis holds
 * geo->nrec[is] traces of real_nt samples (trace-major); writes
 * <tmppath><jobname>seismogram.bin (<jobname>shot<is>_seismogram.bin with
 * per_shot_files) and, with su_output, appends the traces to the SU file
 * (created when first_shot is 1).  SU trace headers (SEG-Y trace
 * identification header, the layout of src/src_cpu/segy.h:27-310): tracl and
 * tracr = trace number in the file from 1, fldr = is + 1, tracf = receiver
 * number from 1, trid = 1, offset = gx - sx, sdepth and gelev = receiver
 * depth as -z, selev = -sz, scalel = scalco = -100 (coordinates in cm),
 * sx, gx, counit = 1, ns = real_nt, dt = real_dt in microseconds.  *ntraces
 * (in, out) counts the traces written to the SU file so far. */
int write_shot_output(int is, const geo2d_t *geo, const float *seis, int real_nt, float real_dt,
                      int first_shot, long long *ntraces);

/* modelling_module (cpumodelling.cpp:158-383) on the GPU with a status
 * return: 0, or the reference's abort code. */
int sam2d_modelling_module(int my_nsrc, geo2d_t *mygeo2d_sp);

#ifdef __cplusplus
}

/* The reference's seam itself: modelling.h:106, same signature, C++ linkage.
 * Runs sam2d_modelling_module; on an error it ends the process with the exit
 * status that the reference's MPI::COMM_WORLD.Abort(code) gives under MPICH,
 * code & 0xff. */
void modelling_module(int my_nsrc, geo2d_t *mygeo2d_sp);
#endif

#endif /* SAM2D_MODELLING_H */
