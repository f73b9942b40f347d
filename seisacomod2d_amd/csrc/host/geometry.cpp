// geometry.cpp -- source/receiver geometry as the reference builds it:
// shot files with split / end-on spreads (src/src_cpu/create_geometry.cpp) and
// explicit sx,sz,gx,gz files (src/src_cpu/read_receiver.cpp).  Coordinates are
// float metres and every expression keeps the reference's float arithmetic,
// including the float-accumulated receiver positions of a spread
// (create_geometry.cpp:336-353); receiver slots the spread loop never writes
// keep (0, 0) as in the reference's calloc'd arrays.  Errors return the
// reference's abort codes.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "sam2d_modelling.h"

#define TRY(x) do { int _r = (x); if (_r) return _r; } while (0)

// set_int / set_float (create_geometry.cpp:149-191): -19 on empty / invalid.
static int g_int(const std::string &t, int &v)
{
    if (t.empty()) return -19;
    try { v = std::stoi(t); } catch (const std::exception &) { return -19; }
    return 0;
}
static int g_float(const std::string &t, float &v)
{
    if (t.empty()) return -19;
    try { v = std::stof(t); } catch (const std::exception &) { return -19; }
    return 0;
}

// Comma-separated fields; more than `maxf` fields is an error (-errc).
static int split_fields(const std::string &line, std::vector<std::string> &tok, int maxf, int errc)
{
    tok.clear();
    std::istringstream stm(line);
    std::string t;
    while (std::getline(stm, t, ',')) {
        tok.push_back(t);
        if ((int)tok.size() > maxf) return errc;
    }
    return (int)tok.size() == maxf ? 0 : errc;
}

// Remainder test of check_geom_para / check_geom_coord: (int)remquof(a, b).
static bool off_grid(float a, float b)
{
    int quo;
    return (int)std::remquo(a, b, &quo) != 0;
}

// check_geom_para (create_geometry.cpp:193-236)
static int check_geom_para(float sx, float sz, float ngeoph)
{
    const int nx = mod_sp->nx, nz = mod_sp->nz;
    const float dx = mod_sp->dx, dz = mod_sp->dz;
    if (off_grid(sx, dx)) return -13;
    if (off_grid(sz, dz)) return -14;
    if (off_grid(ngeoph, dx)) return -15;
    if (sx < 0 || sx > ((nx - 1) * dx)) return -16;
    if (sz < 0 || sz > ((nz - 1) * dz)) return -17;
    return 0;
}

// The spread loop: receivers from start to end every recv_d (float
// accumulation), skipping x == sx when skip_src; at most nrec are written.
static void lay_spread(crd2d_t *rec, int nrec, float start, float end, float recv_d, float sz,
                       float sx, bool skip_src)
{
    float i = start;
    int j = 0;
    while (i <= end && j < nrec) {
        if (!skip_src || i != sx) { rec[j].x = i; rec[j].z = sz; }
        i += recv_d;
        j++;
    }
}

// set_geom2d_split / endonleft / endonright (create_geometry.cpp:265-444)
static int set_receivers(int s, float sx, float sz, int nrec, float nearoff, int type, float recv_d)
{
    const int nx = mod_sp->nx;
    const float xmax = (nx - 1) * mod_sp->dx;
    float start, end;
    if (type == 1) {
        const int tmp1 = (sx - 0.0f) / recv_d;
        const int tmp2 = (xmax - sx) / recv_d;
        if (nearoff == 0.0f) {
            if (tmp1 < nrec / 2) { start = sx - (tmp1 * recv_d); end = start + ((nrec - 1) * recv_d); }
            else if (tmp2 < nrec / 2) { start = sx - ((nrec - tmp2 - 1) * recv_d); end = start + ((nrec - 1) * recv_d); }
            else { start = sx - ((nrec / 2) * recv_d); end = start + ((nrec - 1) * recv_d); }
        } else {
            if (tmp1 < nrec / 2) { start = sx - (tmp1 * recv_d); end = start + ((nrec - 1) * recv_d) + recv_d; }
            else if (tmp2 < nrec / 2) { start = sx - ((nrec - tmp2) * recv_d); end = start + (nrec * recv_d); }
            else {
                start = sx - ((nrec / 2) * recv_d);
                end = start + ((nrec - 1) * recv_d);
                if ((start - recv_d) >= 0) start -= recv_d;
                else end += recv_d;
            }
        }
    } else if (type == 2) {
        end = nearoff == 0.0f ? sx : sx - recv_d;
        start = end - ((nrec - 1) * recv_d);
    } else if (type == 3) {
        start = nearoff == 0.0f ? sx : sx + recv_d;
        end = start + ((nrec - 1) * recv_d);
    } else {
        std::cerr << "\n Error!!!   Wrong geometry type parameter: " << type;
        return -12;
    }
    if (start < 0.0f || end > xmax) {
        std::cerr << "\n Error!!! for source: " << sx << "," << sz << " receivers go out of model";
        return -3;
    }
    lay_spread(geo2d_sp->rec2d_sp[s], nrec, start, end, recv_d, sz, sx, type == 1 && nearoff != 0.0f);
    return 0;
}

extern "C" int read_source2d(const char *shot_file)
{
    // count: lines without "comment" that hold a token (create_geometry.cpp:34-40)
    FILE *fp = std::fopen(shot_file, "r");
    if (!fp) {
        std::cerr << "\nError!!! Unable to open shot file : " << shot_file;
        return -2;
    }
    int nsrc = 0;
    char buf[256], tok0[256];
    while (std::fgets(buf, sizeof buf, fp))
        if (!std::strstr(buf, "comment") && std::sscanf(buf, "%255s", tok0) == 1) nsrc++;
    std::fclose(fp);
    geo2d_sp->nsrc = nsrc;
    geo2d_sp->nrec = (int *)std::calloc(nsrc > 0 ? nsrc : 1, sizeof(int));
    geo2d_sp->src2d_sp = (crd2d_t *)std::calloc(nsrc > 0 ? nsrc : 1, sizeof(crd2d_t));
    geo2d_sp->rec2d_sp = (crd2d_t **)std::calloc(nsrc > 0 ? nsrc : 1, sizeof(crd2d_t *));

    std::ifstream in(shot_file);
    if (!in.is_open()) return -23;
    if (in.peek() == std::ifstream::traits_type::eof()) return -24;
    int default_type = 0;
    std::vector<std::string> t;
    for (int i = 0; i < nsrc; i++) {
        std::string line;
        std::getline(in, line);
        if (line.empty()) return -20;
        TRY(split_fields(line, t, 6, -25));
        float sx, sz, nearoff, ngeoph;
        int nrec, type;
        TRY(g_float(t[0], sx)); TRY(g_float(t[1], sz)); TRY(g_int(t[2], nrec));
        TRY(g_float(t[3], nearoff)); TRY(g_float(t[4], ngeoph)); TRY(g_int(t[5], type));
        if (i == 0) default_type = type;
        if (type != default_type) return -11;
        if (nrec <= 0) return -18;
        geo2d_sp->src2d_sp[i].x = sx;
        geo2d_sp->src2d_sp[i].z = sz;
        geo2d_sp->nrec[i] = nrec;
        geo2d_sp->rec2d_sp[i] = (crd2d_t *)std::calloc(nrec, sizeof(crd2d_t));
        TRY(check_geom_para(sx, sz, ngeoph));
        TRY(set_receivers(i, sx, sz, nrec, nearoff, type, ngeoph));
    }
    return 0;
}

// check_geom_coord (read_receiver.cpp:93-159)
static int check_geom_coord(float sx, float sz, float gx, float gz)
{
    const int nx = mod_sp->nx, nz = mod_sp->nz;
    const float dx = mod_sp->dx, dz = mod_sp->dz;
    if (off_grid(sx, dx)) return -13;
    if (off_grid(sz, dz)) return -14;
    if (off_grid(gx, dx)) return -13;
    if (off_grid(gz, dz)) return -14;
    if (sx < 0 || sx > ((nx - 1) * dx)) return -16;
    if (sz < 0 || sz > ((nz - 1) * dz)) return -17;
    if (gx < 0 || gx > ((nx - 1) * dx)) return -16;
    if (gz < 0 || gz > ((nz - 1) * dz)) return -17;
    return 0;
}

extern "C" int read_recv2d(const char *recv_file)
{
    // count_src (read_receiver.cpp:162-306): shots are runs of equal (sx, sz),
    // sorted by increasing sx.
    std::ifstream in(recv_file);
    if (!in.is_open()) return -3;
    if (in.peek() == std::ifstream::traits_type::eof()) return -4;
    std::vector<std::string> t;
    std::vector<int> counts;
    std::string line;
    std::getline(in, line);
    if (line.empty()) return -21;
    TRY(split_fields(line, t, 4, -26));
    float sx_old, sz_old, gx, gz;
    TRY(g_float(t[0], sx_old)); TRY(g_float(t[1], sz_old)); TRY(g_float(t[2], gx)); TRY(g_float(t[3], gz));
    int nrec = 1;
    while (std::getline(in, line)) {
        if (line.empty()) return -21;
        TRY(split_fields(line, t, 4, -26));
        float sx, sz;
        TRY(g_float(t[0], sx)); TRY(g_float(t[1], sz)); TRY(g_float(t[2], gx)); TRY(g_float(t[3], gz));
        if (sx == sx_old && sz == sz_old) nrec++;
        else if (sx > sx_old) { counts.push_back(nrec); nrec = 1; sx_old = sx; sz_old = sz; }
        else return -22;
    }
    counts.push_back(nrec);
    in.close();

    const int nsrc = (int)counts.size();
    geo2d_sp->nsrc = nsrc;
    geo2d_sp->nrec = (int *)std::calloc(nsrc, sizeof(int));
    geo2d_sp->src2d_sp = (crd2d_t *)std::calloc(nsrc, sizeof(crd2d_t));
    geo2d_sp->rec2d_sp = (crd2d_t **)std::calloc(nsrc, sizeof(crd2d_t *));
    std::ifstream in2(recv_file);
    for (int i = 0; i < nsrc; i++) {
        geo2d_sp->nrec[i] = counts[i];
        geo2d_sp->rec2d_sp[i] = (crd2d_t *)std::calloc(counts[i], sizeof(crd2d_t));
        float sx, sz;
        for (int j = 0; j < counts[i]; j++) {
            std::getline(in2, line);
            split_fields(line, t, 4, -26);
            if (j == 0) { TRY(g_float(t[0], sx)); TRY(g_float(t[1], sz)); }
            TRY(g_float(t[2], gx)); TRY(g_float(t[3], gz));
            TRY(check_geom_coord(sx, sz, gx, gz));
            geo2d_sp->rec2d_sp[i][j].x = gx;
            geo2d_sp->rec2d_sp[i][j].z = gz;
        }
        geo2d_sp->src2d_sp[i].x = sx;
        geo2d_sp->src2d_sp[i].z = sz;
    }
    return 0;
}

// print_single_source_geometry (create_geometry.cpp:446-459): shot 0 as
// "z x" lines, the source first.
extern "C" int write_geometry_dump(const char *path)
{
    FILE *fp = std::fopen(path, "w");
    if (!fp) return -2;
    if (geo2d_sp->nsrc > 0) {
        std::fprintf(fp, "%f %f\n", geo2d_sp->src2d_sp[0].z, geo2d_sp->src2d_sp[0].x);
        for (int j = 0; j < geo2d_sp->nrec[0]; j++)
            std::fprintf(fp, "%f %f\n", geo2d_sp->rec2d_sp[0][j].z, geo2d_sp->rec2d_sp[0][j].x);
    }
    std::fclose(fp);
    return 0;
}
