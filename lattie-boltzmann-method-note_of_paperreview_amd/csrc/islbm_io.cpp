// islbm_io.cpp -- the reference's on-disk formats (fileIO.h, monitor.h) over the C ABI:
// checkpoint/restart of the populations, macroscopic fields and body force, the
// turbulence-statistics sums, and the monitor log.
//
// Padded single-rank layout of the reference (variables.h:16-19 with jp = 1):
// NX6 = nx + 7, NY6 = ny + 7, NZ6 = nz + 6, index J*NX6*NZ6 + K*NX6 + I, unique node (i, j, k)
// at (i + 3, j + 3, k + 3).  Written files carry the periodic images in the x and y ghost
// and duplicate layers (what periodicSW and the single-rank exchange leave) and the
// InitialUsingDftFunc values (rho 1, u 0, f = w_q) in the z ghost layers, which the step
// never writes.
#include "../../include/islbm.h"

#include <sys/stat.h>

#include <cerrno>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

thread_local std::string g_ioerr;

int iofail(int code, const char* fmt, ...)
{
    char buf[768];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_ioerr = buf;
    return code;
}

constexpr int NQ = 19;
const double kW[NQ] = {1.0 / 3.0,  1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0,
                       1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                       1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
// the statistics the reference writes and reads, in its order (fileIO.h:418-451, 475-508)
const char* const kTB[32] = {"U",     "V",     "W",     "P",     "UU",    "UV",    "UW",    "VV",
                             "VW",    "WW",    "PU",    "PV",    "PW",    "KT",    "DUDX2", "DUDY2",
                             "DUDZ2", "DVDX2", "DVDY2", "DVDZ2", "DWDX2", "DWDY2", "DWDZ2", "UUU",
                             "UUV",   "UUW",   "VVU",   "VVV",   "VVW",   "WWU",   "WWV",   "WWW"};

int make_dir(const std::string& d)
{
    if (mkdir(d.c_str(), 0755) != 0 && errno != EEXIST) return iofail(ISLBM_ERR_ARG, "cannot create %s", d.c_str());
    return ISLBM_OK;
}

struct Dims {
    int nx, ny, nz;
    long long NX6, NY6, NZ6, n, npad;
};

Dims dims_of(islbm_solver* s, int* rc)
{
    islbm_params p;
    *rc = islbm_get_params(s, &p);
    Dims d{p.nx, p.ny, p.nz, p.nx + 7, p.ny + 7, p.nz + 6, 0, 0};
    d.n = (long long)p.nx * p.ny * p.nz;
    d.npad = d.NX6 * d.NY6 * d.NZ6;
    return d;
}

int wrapi(long long a, int n)
{
    long long r = a % n;
    return (int)(r < 0 ? r + n : r);
}

// unique [k][j][i] -> padded, z ghosts set to zval
void pad(const Dims& d, const double* uni, double zval, std::vector<double>& out)
{
    out.assign(d.npad, zval);
    for (long long J = 0; J < d.NY6; ++J)
        for (long long K = 3; K < d.NZ6 - 3; ++K)
            for (long long I = 0; I < d.NX6; ++I)
                out[J * d.NX6 * d.NZ6 + K * d.NX6 + I] =
                    uni[((K - 3) * d.ny + wrapi(J - 3, d.ny)) * d.nx + wrapi(I - 3, d.nx)];
}

void unpad(const Dims& d, const std::vector<double>& in, double* uni)
{
    for (int k = 0; k < d.nz; ++k)
        for (int j = 0; j < d.ny; ++j)
            for (int i = 0; i < d.nx; ++i)
                uni[((long long)k * d.ny + j) * d.nx + i] = in[(j + 3) * d.NX6 * d.NZ6 + (k + 3) * d.NX6 + (i + 3)];
}

int write_doubles(const std::string& path, const double* v, size_t n)
{
    FILE* fp = std::fopen(path.c_str(), "wb");
    if (!fp) return iofail(ISLBM_ERR_ARG, "cannot write %s", path.c_str());
    const size_t w = std::fwrite(v, sizeof(double), n, fp);
    std::fclose(fp);
    if (w != n) return iofail(ISLBM_ERR_ARG, "short write %s", path.c_str());
    return ISLBM_OK;
}

int read_doubles(const std::string& path, std::vector<double>& v, size_t n)
{
    FILE* fp = std::fopen(path.c_str(), "rb");
    if (!fp) return iofail(ISLBM_ERR_ARG, "cannot read %s", path.c_str());
    v.assign(n, 0.0);
    const size_t r = std::fread(v.data(), sizeof(double), n, fp);
    const int extra = std::fgetc(fp);
    std::fclose(fp);
    if (r != n || extra != EOF)
        return iofail(ISLBM_ERR_ARG, "%s does not hold %zu doubles (grid mismatch?)", path.c_str(), n);
    return ISLBM_OK;
}

std::string fname(const char* dir, const char* sub, const char* name, int rank)
{
    char b[64];
    std::snprintf(b, sizeof b, "%s_%d.bin", name, rank);
    return std::string(dir) + "/" + sub + "/" + b;
}

}  // namespace

extern "C" {

const char* islbm_io_last_error(void) { return g_ioerr.c_str(); }

// OutputData / fileIO_velocity / fileIO_PDF (fileIO.h:52-67, 262-292): dir/result/*_rank.bin
// and, for rank 0, dir/result/0_force.dat
int islbm_checkpoint_write(islbm_solver* s, const char* dir, int32_t rank)
{
    if (!s || !dir) return iofail(ISLBM_ERR_ARG, "null argument");
    int rc;
    const Dims d = dims_of(s, &rc);
    if (rc) return rc;
    if ((rc = make_dir(dir)) || (rc = make_dir(std::string(dir) + "/result"))) return rc;
    std::vector<double> f((size_t)NQ * d.n), m((size_t)4 * d.n), padded;
    if ((rc = islbm_download_f(s, f.data()))) return rc;
    if ((rc = islbm_download_macros(s, m.data(), m.data() + d.n, m.data() + 2 * d.n, m.data() + 3 * d.n))) return rc;
    const char* mn[4] = {"rho", "u", "v", "w"};
    for (int a = 0; a < 4; ++a) {
        pad(d, m.data() + a * d.n, a == 0 ? 1.0 : 0.0, padded);
        if ((rc = write_doubles(fname(dir, "result", mn[a], rank), padded.data(), padded.size()))) return rc;
    }
    for (int q = 0; q < NQ; ++q) {
        char nm[8];
        std::snprintf(nm, sizeof nm, "f%d", q);
        pad(d, f.data() + q * d.n, kW[q], padded);
        if ((rc = write_doubles(fname(dir, "result", nm, rank), padded.data(), padded.size()))) return rc;
    }
    if (rank == 0) {
        double force = 0.0;
        if ((rc = islbm_get_force(s, &force))) return rc;
        const std::string path = std::string(dir) + "/result/0_force.dat";
        FILE* fp = std::fopen(path.c_str(), "w");
        if (!fp) return iofail(ISLBM_ERR_ARG, "cannot write %s", path.c_str());
        std::fprintf(fp, "%.15lf", force);
        std::fclose(fp);
    }
    return ISLBM_OK;
}

// InitialUsingBkpData (fileIO.h:314-354)
int islbm_checkpoint_read(islbm_solver* s, const char* dir, int32_t rank)
{
    if (!s || !dir) return iofail(ISLBM_ERR_ARG, "null argument");
    int rc;
    const Dims d = dims_of(s, &rc);
    if (rc) return rc;
    const std::string fpath = std::string(dir) + "/result/0_force.dat";
    FILE* fp = std::fopen(fpath.c_str(), "r");
    if (!fp) return iofail(ISLBM_ERR_ARG, "cannot read %s", fpath.c_str());
    double force = 0.0;
    const int got = std::fscanf(fp, "%lf", &force);
    std::fclose(fp);
    if (got != 1) return iofail(ISLBM_ERR_ARG, "%s holds no force value", fpath.c_str());
    std::vector<double> f((size_t)NQ * d.n), m((size_t)4 * d.n), padded;
    const char* mn[4] = {"rho", "u", "v", "w"};
    for (int a = 0; a < 4; ++a) {
        if ((rc = read_doubles(fname(dir, "result", mn[a], rank), padded, d.npad))) return rc;
        unpad(d, padded, m.data() + a * d.n);
    }
    for (int q = 0; q < NQ; ++q) {
        char nm[8];
        std::snprintf(nm, sizeof nm, "f%d", q);
        if ((rc = read_doubles(fname(dir, "result", nm, rank), padded, d.npad))) return rc;
        unpad(d, padded, f.data() + q * d.n);
    }
    if ((rc = islbm_upload_f(s, f.data()))) return rc;
    if ((rc = islbm_upload_macros(s, m.data(), m.data() + d.n, m.data() + 2 * d.n, m.data() + 3 * d.n))) return rc;
    return islbm_set_force(s, force);
}

// Launch_OutputTB / OutputTBData (fileIO.h:379-451): dir/statistics/NAME/NAME_rank.bin over
// I = 3 .. NX6-4, J = 3 .. NY6-4, K = 3 .. NZ6-4 (k outer, i inner) and dir/statistics/accu.dat
int islbm_stats_write(islbm_solver* s, const char* dir, int32_t rank)
{
    if (!s || !dir) return iofail(ISLBM_ERR_ARG, "null argument");
    int rc;
    const Dims d = dims_of(s, &rc);
    if (rc) return rc;
    const std::string sd = std::string(dir) + "/statistics";
    if ((rc = make_dir(dir)) || (rc = make_dir(sd))) return rc;
    int64_t count = 0;
    if ((rc = islbm_stats_count(s, &count))) return rc;
    if (rank == 0) {
        FILE* fp = std::fopen((sd + "/accu.dat").c_str(), "w");
        if (!fp) return iofail(ISLBM_ERR_ARG, "cannot write %s/accu.dat", sd.c_str());
        std::fprintf(fp, "%d", (int)count);
        std::fclose(fp);
    }
    std::vector<double> uni(d.n), out;
    out.reserve((size_t)(d.nx + 1) * (d.ny + 1) * d.nz);
    for (const char* nm : kTB) {
        if ((rc = islbm_stats_download(s, nm, uni.data()))) return rc;
        if ((rc = make_dir(sd + "/" + nm))) return rc;
        out.clear();
        for (int K = 3; K < d.NZ6 - 3; ++K)
            for (int J = 3; J < d.NY6 - 3; ++J)
                for (int I = 3; I < d.NX6 - 3; ++I)
                    out.push_back(uni[((long long)(K - 3) * d.ny + wrapi(J - 3, d.ny)) * d.nx + wrapi(I - 3, d.nx)]);
        if ((rc = write_doubles(fname(sd.c_str(), nm, nm, rank), out.data(), out.size()))) return rc;
    }
    return ISLBM_OK;
}

// InitialTBUsingBkpData (fileIO.h:453-508); PP is not part of the files and restarts at 0
int islbm_stats_read(islbm_solver* s, const char* dir, int32_t rank)
{
    if (!s || !dir) return iofail(ISLBM_ERR_ARG, "null argument");
    int rc;
    const Dims d = dims_of(s, &rc);
    if (rc) return rc;
    const std::string sd = std::string(dir) + "/statistics";
    FILE* fp = std::fopen((sd + "/accu.dat").c_str(), "r");
    if (!fp) return iofail(ISLBM_ERR_ARG, "cannot read %s/accu.dat", sd.c_str());
    int count = 0;
    const int got = std::fscanf(fp, "%d", &count);
    std::fclose(fp);
    if (got != 1) return iofail(ISLBM_ERR_ARG, "accu.dat holds no count");
    if ((rc = islbm_stats_reset(s))) return rc;
    const size_t nf = (size_t)(d.nx + 1) * (d.ny + 1) * d.nz;
    std::vector<double> in, uni(d.n);
    for (const char* nm : kTB) {
        if ((rc = read_doubles(fname(sd.c_str(), nm, nm, rank), in, nf))) return rc;
        size_t t = 0;
        for (int K = 3; K < d.NZ6 - 3; ++K)
            for (int J = 3; J < d.NY6 - 3; ++J)
                for (int I = 3; I < d.NX6 - 3; ++I, ++t)
                    if (I - 3 < d.nx && J - 3 < d.ny)
                        uni[((long long)(K - 3) * d.ny + (J - 3)) * d.nx + (I - 3)] = in[t];
        if ((rc = islbm_stats_upload(s, nm, uni.data()))) return rc;
    }
    return islbm_stats_set_count(s, count);
}

// Launch_Monitor (monitor.h:4-26): "step\t v\t u\t w\t force" appended to path
int islbm_monitor_append(islbm_solver* s, const char* path)
{
    if (!s || !path) return iofail(ISLBM_ERR_ARG, "null argument");
    double m[4];
    int64_t step = 0;
    int rc;
    if ((rc = islbm_monitor(s, m)) || (rc = islbm_step_count(s, &step))) return rc;
    FILE* fp = std::fopen(path, "a");
    if (!fp) return iofail(ISLBM_ERR_ARG, "cannot append to %s", path);
    std::fprintf(fp, "%d\t %.15lf\t %.15lf\t %.15lf\t %.15lf\n", (int)step, m[0], m[1], m[2], m[3]);
    std::fclose(fp);
    return ISLBM_OK;
}

}  // extern "C"
