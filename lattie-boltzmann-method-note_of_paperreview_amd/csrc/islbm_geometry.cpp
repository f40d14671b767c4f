// islbm_geometry.cpp -- FP64 host geometry of one periodic-hill solve in unique-node form.
//
// Restates the reference's set-up path for a single rank (jp = 1) without ghost layers:
// GenerateMesh_X/Y/Z (initialization.h:51-183), GetNonuniParameter (initializationTool.h:16-39),
// GetIntrplParameter_X/Y (initialization.h:203-229) and BFLInitialization
// (initialization.h:260-360).  The per-node xi weights (GetIntrplParameter_Xi,
// initialization.h:231-258) are computed on the device in FP64 (islbm_kernels.hip).
// Unique node (i, j, k) is the reference's padded node (i+3, j+3, k+3).
#include "islbm_internal.h"

#include <cmath>
#include <cstdio>

namespace islbm {

namespace {

constexpr int kBisectCap = 400;  // the reference's bisections have no cap

// tanhFunction: initializationTool.h:4-7
double tanh_map(double L, double lattice, double a, int j, int n)
{
    return L / 2.0 + lattice / 2.0 +
           ((L / 2.0) / a) * std::tanh((-1.0 + 2.0 * (double)(j) / (double)(n)) / 2.0 * std::log((1.0 + a) / (1.0 - a)));
}

// GetNonuniParameter: bisect a in [0.1, 1] until the first cell equals minSize to 1e-14.
double bisect_a(double lz, double minsize, int nz, double hill0)
{
    const double total = lz - hill0 - minsize;
    double lo = 0.1, hi = 1.0, mid = 0.0, d;
    int it = 0;
    do {
        mid = (lo + hi) / 2.0;
        d = tanh_map(total, minsize, mid, 1, nz - 1) - tanh_map(total, minsize, mid, 0, nz - 1);
        if (d - minsize >= 0.0)
            lo = mid;
        else
            hi = mid;
    } while (std::fabs(d - minsize) > 1e-14 && ++it < kBisectCap);
    return mid;
}

// GetDeltaHorizontal: initializationTool.h:90-107
double delta_horizontal(double zt, double y_in, double y_out, double y_pt, double ly, double al, double ms)
{
    double t0 = y_in, t1 = y_out, mid;
    int it = 0;
    do {
        mid = (t0 + t1) / 2.0;
        if (hill_profile(mid, ly, al, 0) >= zt)
            t0 = mid;
        else
            t1 = mid;
    } while (std::fabs(hill_profile(mid, ly, al, 0) - zt) > 1e-13 && ++it < kBisectCap);
    return ms - 2.0 * std::fabs(y_pt - mid);
}

// GetDelta45Degree: initializationTool.h:125-154
double delta_45(double zp, double yp, double y_zdom, double y_ydom, double ly, double al, double ms)
{
    double t0 = y_zdom, t1 = y_ydom, mid;
    int it = 0;
    do {
        mid = (t0 + t1) / 2.0;
        if (std::fabs(hill_profile(mid, ly, al, 0) - zp) > std::fabs(mid - yp))
            t0 = mid;
        else
            t1 = mid;
    } while (std::fabs(std::fabs(hill_profile(mid, ly, al, 0) - zp) - std::fabs(mid - yp)) > 1e-13 &&
             ++it < kBisectCap);
    return ms - 2.0 * std::fabs(yp - mid);
}

void lagrange7(double* w7, double x, const double* nodes7)
{
    for (int m = 0; m < 7; ++m) w7[m] = lagrange7_weight(x, nodes7, m);
}

}  // namespace

std::string build_host_geometry(const islbm_params& p, HostGeom& g)
{
    char msg[256];
    if (p.nx < 4 || p.ny < 8 || p.nz < 8) {
        std::snprintf(msg, sizeof msg, "grid %d x %d x %d too small (need nx>=4, ny>=8, nz>=8)", p.nx, p.ny, p.nz);
        return msg;
    }
    if (!(p.lx > 0 && p.ly > 0 && p.lz > 1.0 && p.cfl > 0 && p.lxi > 0 && p.hill_alpha > 0))
        return "domain lengths, cfl and lxi must be positive and lz > 1 (hill height)";
    if (!(p.re > 0)) return "re must be positive";
    if (p.tau <= 0 && !(p.u_ref > 0)) return "either tau > 0 or u_ref > 0 is required";
    if (p.ndtfrc <= 0) return "ndtfrc must be positive";
    if (p.cs_smag < 0) return "cs_smag must be >= 0";
    if (p.mesh_mode != ISLBM_MESH_BISECT && p.mesh_mode != ISLBM_MESH_STRETCH) return "unknown mesh_mode";
    if (p.mesh_mode == ISLBM_MESH_STRETCH && !(p.stretch > 0)) return "stretch must be positive";
    if (p.xi_near_wall != ISLBM_XI_REFERENCE && p.xi_near_wall != ISLBM_XI_CLAMPED_LINEAR)
        return "unknown xi_near_wall";

    const int nx = p.nx, ny = p.ny, nz = p.nz;
    g.nx = nx; g.ny = ny; g.nz = nz;
    // variables.h:23,76,95-99
    g.minsize = (p.lz - 1.0) / nz * p.cfl;
    g.dt = g.minsize;
    if (p.tau > 0) {
        g.tau = p.tau;
        g.nu = (g.tau - 0.5) / 3.0 * g.dt;
    } else {
        g.nu = p.u_ref / p.re;
        g.tau = 3.0 * g.nu / g.dt + 0.5;
    }
    g.uref = p.re * g.nu;
    if (!(g.tau > 0.5)) return "relaxation time must exceed 1/2";
    g.force0 = (8.0 * g.nu * g.uref) / (p.lz * p.lz) * 5.0;  // initialization.h:46
    const double ms = g.minsize;
    const double ly = p.ly, al = p.hill_alpha;
    const int flat = p.flat;

    // GenerateMesh_X/Y: uniform, x_h[I] = dx*(I-3)
    g.dx = p.lx / (double)nx;
    g.dy = p.ly / (double)ny;
    auto xat = [&](int ip) { return g.dx * ((double)ip); };  // ip = unique index, may be outside [0,nx)
    auto yat = [&](int jp) { return g.dy * ((double)jp); };
    g.x.resize(nx); g.y.resize(ny);
    for (int i = 0; i < nx; ++i) g.x[i] = xat(i);
    for (int j = 0; j < ny; ++j) g.y[j] = yat(j);

    // GenerateMesh_Z
    const double hill0 = hill_profile(0.0, ly, al, flat);
    g.a = (p.mesh_mode == ISLBM_MESH_BISECT) ? bisect_a(p.lz, ms, nz, hill0) : std::tanh(p.stretch / 2.0);
    if (!(g.a > 0 && g.a < 1)) return "tanh mesh parameter outside (0,1)";
    g.zc.resize((size_t)ny * nz);
    for (int j = 0; j < ny; ++j) {
        const double H = hill_profile(g.y[j], ly, al, flat);
        const double total = p.lz - H - ms;
        for (int k = 0; k < nz; ++k) g.zc[(size_t)j * nz + k] = tanh_map(total, ms, g.a, k, nz - 1) + H;
    }
    g.xi.resize(nz);
    for (int k = 0; k < nz; ++k) g.xi[k] = tanh_map(p.lxi, ms, g.a, k, nz - 1) - ms / 2.0;

    // GetIntrplParameter_X/Y: departure x -+ minSize on nodes i-3..i+3
    double nodes[7], w7[7];
    g.xw0.resize((size_t)nx * 7); g.xw2.resize((size_t)nx * 7);
    for (int i = 0; i < nx; ++i) {
        for (int m = 0; m < 7; ++m) nodes[m] = xat(i - 3 + m);
        lagrange7(&g.xw0[(size_t)i * 7], xat(i) - ms, nodes);
        lagrange7(&g.xw2[(size_t)i * 7], xat(i) + ms, nodes);
    }
    g.yw0.resize((size_t)ny * 7); g.yw2.resize((size_t)ny * 7);
    for (int j = 0; j < ny; ++j) {
        for (int m = 0; m < 7; ++m) nodes[m] = yat(j - 3 + m);
        lagrange7(&g.yw0[(size_t)j * 7], yat(j) - ms, nodes);
        lagrange7(&g.yw2[(size_t)j * 7], yat(j) + ms, nodes);
    }

    // GetXiParameter for the BFL layers (k = 0, 1: stencil start 0)
    auto xi_w = [&](double* out7, double pz, double py) {
        const double H = hill_profile(py, ly, al, flat);
        const double L = p.lz - H - ms;
        const double pxi = p.lxi * (pz - (H + ms / 2.0)) / L;
        if (p.xi_near_wall)
            xi_weights7(pxi, g.xi.data(), 0, nz, 1, out7);  // layers 0 and 1 share stencil 0
        else
            lagrange7(out7, pxi, g.xi.data());
    };

    // BFLInitialization: flags, q and weights on the first two layers
    const size_t nb = 2 * (size_t)ny;
    g.bflag.assign(NBFL * nb, 0);
    g.bq.assign(NBFL * nb, 0.0);
    g.bw.assign((size_t)NBW * nb * 7, 0.0);
    auto bwp = [&](int which, size_t id) { return &g.bw[((size_t)which * nb + id) * 7]; };
    if (!flat) {
        const int NX6 = nx + 7;
        const int ic = NX6 / 2 - 3;  // unique index of the reference's x_h[NX6/2]
        double xn[7];
        for (int m = 0; m < 7; ++m) xn[m] = xat(ic - 3 + m);
        const double xm = xat(ic);
        for (int k = 0; k < 2; ++k)
            for (int j = 0; j < ny; ++j) {
                const size_t id = (size_t)k * ny + j;
                const double zz = g.zc[(size_t)j * nz + k], yy = g.y[j];
                // IsBFLBCNeeded: initializationTool.h:81-88
                g.bflag[B3 * nb + id] = hill_profile(yy - ms, ly, al, 0) > zz;
                g.bflag[B4 * nb + id] = hill_profile(yy + ms, ly, al, 0) > zz;
                g.bflag[B15 * nb + id] = hill_profile(yy - ms, ly, al, 0) > zz - ms;
                g.bflag[B16 * nb + id] = hill_profile(yy + ms, ly, al, 0) > zz - ms;
                double yn[7];
                for (int m = 0; m < 7; ++m) yn[m] = yat(j - 3 + m);
                double delta;
                if (g.bflag[B3 * nb + id]) {
                    delta = delta_horizontal(zz, yy - ms, yy, yy, ly, al, ms);
                    g.bq[B3 * nb + id] = (ms - delta) / 2.0 / ms;
                    lagrange7(bwp(BX38, id), xm - delta, xn);
                    lagrange7(bwp(BX37, id), xm + delta, xn);
                    lagrange7(bwp(BY378, id), yy + delta, yn);
                    xi_w(bwp(BXI378, id), zz, yy + delta);
                }
                if (g.bflag[B4 * nb + id]) {
                    delta = delta_horizontal(zz, yy + ms, yy, yy, ly, al, ms);
                    g.bq[B4 * nb + id] = (ms - delta) / 2.0 / ms;
                    lagrange7(bwp(BX49, id), xm + delta, xn);
                    lagrange7(bwp(BX410, id), xm - delta, xn);
                    lagrange7(bwp(BY4910, id), yy - delta, yn);
                    xi_w(bwp(BXI4910, id), zz, yy - delta);
                }
                if (g.bflag[B15 * nb + id]) {
                    delta = delta_45(zz, yy, yy, yy - ms, ly, al, ms);
                    g.bq[B15 * nb + id] = (ms - delta) / 2.0 / ms;
                    lagrange7(bwp(BY15, id), yy + delta, yn);
                    xi_w(bwp(BXI15, id), zz + delta, yy + delta);
                }
                if (g.bflag[B16 * nb + id]) {
                    delta = delta_45(zz, yy, yy, yy + ms, ly, al, ms);
                    g.bq[B16 * nb + id] = (ms - delta) / 2.0 / ms;
                    lagrange7(bwp(BY16, id), yy - delta, yn);
                    xi_w(bwp(BXI16, id), zz + delta, yy - delta);
                }
            }
    }

    // AccumulateUbulk weights (evolution.h:826-840) on the periodic image plane of the
    // reference's ghost plane J = 0, i.e. y = -3 dy (wrapped to LY - 3 dy).
    {
        g.flux_plane = wrap_index(ny - 3, ny);
        const double yg = g.dy * ((double)(-3));
        const double H = hill_profile(yg, ly, al, flat);
        const double total = p.lz - H - ms;
        std::vector<double> zg(nz + 2);
        zg[0] = H;
        for (int k = 0; k < nz; ++k) zg[k + 1] = tanh_map(total, ms, g.a, k, nz - 1) + H;
        zg[nz + 1] = p.lz;
        g.fluxw.resize((size_t)nz * nx);
        g.flux_dx.resize(nx);
        g.flux_dz.resize(nz);
        for (int i = 0; i < nx; ++i) g.flux_dx[i] = (xat(i + 1) - xat(i - 1)) / 2.0;
        for (int k = 0; k < nz; ++k) {
            const double dzl = (zg[k + 2] - zg[k]) / 2.0;
            g.flux_dz[k] = dzl;
            for (int i = 0; i < nx; ++i) g.fluxw[(size_t)k * nx + i] = g.flux_dx[i] * dzl;
        }
    }

    // Smagorinsky filter width (dx dy dz_local)^(1/3) per (k, j)
    g.les_delta.resize((size_t)nz * ny);
    for (int j = 0; j < ny; ++j) {
        const double H = hill_profile(g.y[j], ly, al, flat);
        for (int k = 0; k < nz; ++k) {
            const double zm = k == 0 ? H : g.zc[(size_t)j * nz + k - 1];
            const double zp = k == nz - 1 ? p.lz : g.zc[(size_t)j * nz + k + 1];
            g.les_delta[(size_t)k * ny + j] = std::cbrt(g.dx * g.dy * ((zp - zm) / 2.0));
        }
    }
    return "";
}

std::vector<double> zderiv_weights(const HostGeom& g)
{
    const int ny = g.ny, nz = g.nz;
    std::vector<double> w((size_t)ny * nz * 5, 0.0);
    for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nz; ++k) {
            int n = k - 2;
            n = n < 0 ? 0 : (n > nz - 5 ? nz - 5 : n);
            const double* zl = &g.zc[(size_t)j * nz + n];
            const double x = g.zc[(size_t)j * nz + k];
            for (int m = 0; m < 5; ++m) {
                // L_m'(x) = sum_{l != m} 1/(z_m - z_l) prod_{r != m, l} (x - z_r)/(z_m - z_r)
                double d = 0.0;
                for (int l = 0; l < 5; ++l) {
                    if (l == m) continue;
                    double t = 1.0 / (zl[m] - zl[l]);
                    for (int r = 0; r < 5; ++r)
                        if (r != m && r != l) t *= (x - zl[r]) / (zl[m] - zl[r]);
                    d += t;
                }
                w[((size_t)j * nz + k) * 5 + m] = d;
            }
        }
    return w;
}

}  // namespace islbm
