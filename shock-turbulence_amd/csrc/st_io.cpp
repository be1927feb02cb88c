// On-disk formats on either side of the pipeline: the restart file of
// Write_Restart / Read_Restart (src/out.f:493-692) and the input deck of
// Define_Parameters (src/define_periodic_sf.f:20-80).  Host code only.
#include <cctype>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/st_api.h"

namespace {

// The restart header: the global extents (Ny, Nz, Nx, 5) as four default
// integers, the time as REAL*8 and the step as a default integer, written
// back to back by rank 0 (src/out.f:538-545); the data follow at byte
// 5*4 + 8 = 28 (:560) as Q(Ny,Nz,Nx,5) in Fortran order (:562-568).
constexpr long HEADER_BYTES = 28;

int write_all(FILE *f, const void *p, size_t n) { return fwrite(p, 1, n, f) == n ? 0 : -1; }
int read_all(FILE *f, void *p, size_t n) { return fread(p, 1, n, f) == n ? 0 : -1; }

}  // namespace

extern "C" int st_restart_write(const char *path, int nx, int ny, int nz, const double *Q, double time, int step)
{
    if (!path || !Q || nx < 1 || ny < 1 || nz < 1) return -1;
    FILE *f = fopen(path, "wb");
    if (!f) return -2;
    const int32_t g[4] = {ny, nz, nx, 5};
    const int32_t st = step;
    const size_t n = (size_t)nx * ny * nz * 5;
    int rc = write_all(f, g, sizeof(g));
    if (!rc) rc = write_all(f, &time, sizeof(time));
    if (!rc) rc = write_all(f, &st, sizeof(st));
    if (!rc) rc = write_all(f, Q, n * sizeof(double));
    if (fclose(f) || rc) return -3;
    return 0;
}

extern "C" int st_restart_read(const char *path, int nx, int ny, int nz, double *Q, double *time, int *step)
{
    if (!path || !Q || nx < 1 || ny < 1 || nz < 1) return -1;
    FILE *f = fopen(path, "rb");
    if (!f) return -2;
    int32_t g[4], st = 0;
    double t = 0.0;
    int rc = read_all(f, g, sizeof(g));
    if (!rc) rc = read_all(f, &t, sizeof(t));
    if (!rc) rc = read_all(f, &st, sizeof(st));
    if (rc) {
        fclose(f);
        return -3;
    }
    if (g[0] != ny || g[1] != nz || g[2] != nx || g[3] != 5) {
        fclose(f);
        return -4;                               // extents of another grid
    }
    if (fseek(f, HEADER_BYTES, SEEK_SET)) rc = -3;
    if (!rc) rc = read_all(f, Q, (size_t)nx * ny * nz * 5 * sizeof(double)) ? -3 : 0;
    fclose(f);
    if (rc) return rc;
    if (time) *time = t;
    if (step) *step = st;
    return 0;
}

namespace {

// One record of the deck.  List-directed reads take the first token
// (Fortran's d exponent read as e); paths are the first 20 characters (A20).
struct Deck {
    FILE *f;
    char line[512];
    int ok = 1;

    bool next()
    {
        if (!fgets(line, sizeof(line), f)) {
            ok = 0;
            line[0] = 0;
            return false;
        }
        return true;
    }
    std::string token()
    {
        next();
        const char *p = line;
        while (*p && isspace((unsigned char)*p)) p++;
        std::string t;
        while (*p && !isspace((unsigned char)*p) && *p != ',') t += *p++;
        for (auto &ch : t)
            if (ch == 'd' || ch == 'D') ch = 'e';
        if (t.empty()) ok = 0;
        return t;
    }
    double real()
    {
        const std::string t = token();
        char *end = nullptr;
        const double v = strtod(t.c_str(), &end);
        if (!end || *end) ok = 0;
        return v;
    }
    int integer()
    {
        const std::string t = token();
        char *end = nullptr;
        const long v = strtol(t.c_str(), &end, 10);
        if (!end || *end) ok = 0;
        return (int)v;
    }
    void a20(char *out)
    {
        next();
        char buf[21];
        std::strncpy(buf, line, 20);
        buf[20] = 0;
        int n = (int)std::strlen(buf);
        while (n > 0 && (isspace((unsigned char)buf[n - 1]) || buf[n - 1] == '\n')) buf[--n] = 0;
        const char *p = buf;
        while (*p == ' ') p++;
        std::strcpy(out, p);
    }
};

}  // namespace

extern "C" int st_deck_read(const char *path, st_deck *d)
{
    if (!path || !d) return -1;
    FILE *f = fopen(path, "r");
    if (!f) return -2;
    std::memset(d, 0, sizeof(*d));
    Deck r{f};
    r.next();                                    // comment (A80)
    r.a20(d->vis_path);
    r.a20(d->restart_path);
    r.a20(d->stats_path);
    r.a20(d->probe_path);
    r.a20(d->source_path);
    r.a20(d->prefix);
    r.next();
    r.next();                                    // grid parameters
    d->grid_x = r.integer();
    d->grid_y = r.integer();
    d->x_start = r.real();
    d->y_start = r.real();
    d->z_start = r.real();
    d->period_x = r.real();
    d->period_y = r.real();
    d->period_z = r.real();
    r.next();
    r.next();                                    // flow parameters
    d->Re = r.real();
    d->Pr = r.real();
    d->Sc = r.real();
    d->gamma = r.real();
    d->vis_exp = r.real();
    d->bulk_ratio = r.real();
    d->T_ratio = r.real();
    d->hyper_param = r.integer();
    d->part_param = r.integer();
    r.next();
    r.next();                                    // time advancement parameters
    d->march_mode = r.integer();
    d->Tstep = r.real();
    d->cfl_cal = r.real();
    d->Nsteps = r.integer();
    d->run_time = r.real();
    d->comp_stats = r.integer();
    r.next();
    r.next();                                    // input/output parameters
    d->out_grid = r.integer();
    d->out_init = r.integer();
    d->out_incr_x = r.integer();
    d->out_incr_y = r.integer();
    d->out_incr_z = r.integer();
    d->out_restart = r.integer();
    d->out_vis = r.integer();
    d->out_probe = r.integer();
    d->out_stats = r.integer();
    d->out_source = r.integer();
    fclose(f);
    return r.ok ? 0 : -3;
}
