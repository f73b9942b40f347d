// jobcard.cpp -- the reference's job-card front end (src/src_cpu/json_parser.cpp)
// over the process globals of include/sam2d_modelling.h.  Errors return the
// reference's abort code instead of calling MPI::COMM_WORLD.Abort.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <map>
#include <stdexcept>
#include <string>

#include "sam2d_modelling.h"

static mod_t g_mod;
static wave_t g_wave;
static geo2d_t g_geo;
static job_t g_job;
mod_t *mod_sp = &g_mod;
wave_t *wave_sp = &g_wave;
geo2d_t *geo2d_sp = &g_geo;
job_t *job_sp = &g_job;
sam2d_run_options sam2d_opts = {32, 0, 1, 0, 0};

// json_parser.cpp:13-37: the card's file name without directory and
// extension, followed by "_".
extern "C" void set_job_name(const char *job)
{
    std::string s(job);
    const size_t slash = s.rfind('/');
    std::string base = slash == std::string::npos ? s : s.substr(slash + 1);
    const size_t dot = base.rfind('.');
    if (dot != std::string::npos) base = base.substr(0, dot);
    std::snprintf(job_sp->jobname, sizeof job_sp->jobname, "%s_", base.c_str());
}

static int missing(const char *what)
{
    std::cout << "\n Error!!! missing " << what << " information in json file....\n";
    return -6;
}

// str_to_int / str_to_float (json_parser.cpp:92-120): std::stoi / std::stof
// of a non-empty value, prefix parse, error on no digits.
static int to_int(const std::string &v, int &out, const char *what)
{
    if (v.empty()) return missing(what);
    try { out = std::stoi(v); } catch (const std::exception &) { return missing(what); }
    return 0;
}
static int to_float(const std::string &v, float &out, const char *what)
{
    if (v.empty()) return missing(what);
    try { out = std::stof(v); } catch (const std::exception &) { return missing(what); }
    return 0;
}

#define TRY(x) do { int _r = (x); if (_r) return _r; } while (0)

extern "C" int read_job_card(const char *job)
{
    // read_json_objects (json_parser.cpp:40-83): lines holding a quote and a
    // colon and neither "comment" nor "Comment"; "key" : "value"; the first
    // occurrence of a key wins.
    FILE *fp = std::fopen(job, "r");
    if (!fp) {
        std::cerr << "\n Error!!! Unable to open input job card : " << job;
        return -2;
    }
    std::map<std::string, std::string> kv;
    char line[256];
    while (std::fgets(line, sizeof line, fp)) {
        if (std::strstr(line, "\"") && std::strstr(line, ":") && !std::strstr(line, "comment") &&
            !std::strstr(line, "Comment")) {
            char key[32] = "", value[224] = "";
            std::sscanf(line, " \"%31[^\"]\" : \"%223[^\"]\" ", key, value);
            kv.insert(std::make_pair(std::string(key), std::string(value)));
        }
    }
    std::fclose(fp);

    // set_json_object (json_parser.cpp:122-300)
    auto get = [&](const char *k, std::string &out) -> bool {
        auto it = kv.find(k);
        if (it == kv.end()) return false;
        out = it->second;
        return true;
    };
    std::string v;
    if (!get("Velocity", v)) return missing("velocity file name");
    std::snprintf(mod_sp->velfile, sizeof mod_sp->velfile, "%s", v.c_str());
    if (!get("Density", v)) return missing("density file name");
    std::snprintf(mod_sp->densfile, sizeof mod_sp->densfile, "%s", v.c_str());
    if (!get("NX", v)) return missing("NX (number of grid points in X direction)");
    TRY(to_int(v, mod_sp->nx, "NX (number of grid points in X direction)"));
    if (!get("NZ", v)) return missing("NZ (number of grid points in Z direction)");
    TRY(to_int(v, mod_sp->nz, "NZ (number of grid points in Z direction)"));
    if (!get("DX", v)) return missing("grid spacing in X direction");
    TRY(to_float(v, mod_sp->dx, "grid spacing in X direction"));
    if (!get("DZ", v)) return missing("grid spacing in Z direction");
    TRY(to_float(v, mod_sp->dz, "grid spacing in Z direction"));
    if (!get("Time", v)) return missing("record time");
    TRY(to_float(v, wave_sp->rec_len, "record time"));
    if (!get("FD DT", v)) return missing("FD sampling interval (DT)");
    TRY(to_float(v, wave_sp->fd_dt_sec, "FD sampling interval (DT)"));
    if (!get("DATA DT", v)) return missing("Seismogram sampling interval (DT)");
    TRY(to_float(v, wave_sp->real_dt_sec, "Seismogram sampling interval (DT)"));
    if (std::fmod(wave_sp->real_dt_sec, wave_sp->fd_dt_sec) != 0)
        return missing("DATA DT must be multiple of FD DT (DATA_DT>=FD_DT & DATA_DT%FD_DT=0)");
    if (!get("Frequency", v)) return missing("dominant frequency");
    TRY(to_float(v, wave_sp->dom_freq, "dominant frequency"));
    if (!get("Read source receiver flag", v)) return missing("read source-receiver flag");
    TRY(to_int(v, job_sp->geomflag, "read source-receiver flag"));
    if (job_sp->geomflag < 0 || job_sp->geomflag > 1)
        return missing("wrong flag value for read source-receiver flag");
    if (job_sp->geomflag == 0) {
        if (!get("if choice=0, Geometry file", v)) return missing("source-receiver geometry file name");
        std::snprintf(job_sp->geomfile, sizeof job_sp->geomfile, "%s", v.c_str());
    } else {
        if (!get("if choice=1, Shot file", v)) return missing("shot file name");
        std::snprintf(job_sp->shotfile, sizeof job_sp->shotfile, "%s", v.c_str());
    }
    if (!get("Output Directory", v)) return missing("Output directory path");
    if (v.empty() || v.back() != '/') v += "/";
    std::snprintf(job_sp->tmppath, sizeof job_sp->tmppath, "%s", v.c_str());
    if (!get("Free surface", v)) return missing("free surface flag");
    TRY(to_int(v, job_sp->frsf, "free surface flag"));
    if (job_sp->frsf < 0 || job_sp->frsf > 1) return missing("wrong flag value for free surface choice");
    if (!get("Job type New or Restart", v)) return missing("job type New or Restart");
    std::snprintf(job_sp->jbtype, sizeof job_sp->jbtype, "%s", v.c_str());
    if (std::strcmp(job_sp->jbtype, "New") != 0 && std::strcmp(job_sp->jbtype, "Restart") != 0) {
        std::cerr << "\n Error!!! wrong value is specified for job status";
        return -6;
    }
    if (!get("Order of FD Operator", v)) return missing("order of FD Operator");
    TRY(to_int(v, job_sp->fdop, "order of FD Operator"));
    job_sp->adj = 0;
    return 0;
}
