// ref_geom_driver.cpp -- test harness that plays the role of the reference's absent main.cu
// for the host-only geometry and MRT-table headers, and prints their outputs as JSON.
//
// TEST INFRASTRUCTURE ONLY: built by oracle/Makefile into oracle/_ref/ straight from the
// reference sources under /root/reference (variables.h, model.h, initializationTool.h,
// MRT_Matrix.h); nothing from those files is copied here.  The compile-time parameter set
// of variables.h is the one it ships (NX 32, NY 128, NZ 64, jp 4, tau 0.6833, Re 50).
// The translation unit supplies what a CUDA main.cu would: <cmath>, <cstdio> and the
// standard min/max (model.h:17,33 call min/max on doubles).
#include <algorithm>
#include <cmath>
#include <cstdio>
using std::max;
using std::min;

#include "variables.h"
#include "model.h"
#include "initializationTool.h"
#include "MRT_Matrix.h"

static void arr(const char* name, const double* v, int n, bool last = false)
{
    std::printf("\"%s\": [", name);
    for (int i = 0; i < n; ++i) std::printf("%s%.17g", i ? ", " : "", v[i]);
    std::printf("]%s\n", last ? "" : ",");
}

int main()
{
    std::printf("{\n");
    std::printf("\"NX\": %d, \"NY\": %d, \"NZ\": %d, \"NZ6\": %d, \"jp\": %d,\n", NX, NY, NZ, NZ6, jp);
    std::printf("\"LX\": %.17g, \"LY\": %.17g, \"LZ\": %.17g, \"LXi\": %.17g, \"CFL\": %.17g,\n",
                (double)LX, (double)LY, (double)LZ, (double)LXi, (double)CFL);
    std::printf("\"minSize\": %.17g, \"dt\": %.17g, \"tau\": %.17g, \"niu\": %.17g, \"Uref\": %.17g, \"Re\": %d,\n",
                (double)minSize, (double)dt, (double)tau, (double)niu, (double)Uref, (int)Re);

    // HillFunction (model.h:4-68) on a grid that covers both wraps and every breakpoint.
    const int nh = 2401;
    static double hy[2401], hv[2401];
    for (int i = 0; i < nh; ++i) {
        hy[i] = -1.5 + 12.0 * (double)i / (double)(nh - 1);
        hv[i] = HillFunction(hy[i]);
    }
    arr("hill_y", hy, nh);
    arr("hill_h", hv, nh);
    // breakpoints in y (units of h), from both sides
    const double bps[6] = {9. / 28., 14. / 28., 20. / 28., 30. / 28., 40. / 28., 54. / 28.};
    double by[24], bv[24];
    int nb = 0;
    for (int i = 0; i < 6; ++i) {
        double b = (54. / 28.) * (bps[i] * 28. / 54.);
        by[nb] = b; bv[nb++] = HillFunction(b);
        by[nb] = std::nextafter(b, 10.0); bv[nb] = HillFunction(by[nb]); nb++;
        by[nb] = LY - b; bv[nb] = HillFunction(by[nb]); nb++;
        by[nb] = std::nextafter(LY - b, -10.0); bv[nb] = HillFunction(by[nb]); nb++;
    }
    arr("hill_break_y", by, nb);
    arr("hill_break_h", bv, nb);

    // GetNonuniParameter (initializationTool.h:16-39) and tanhFunction (:4-7)
    double a = GetNonuniParameter();
    std::printf("\"a_bisect\": %.17g,\n", a);
    double tz[NZ6];
    const double total0 = LZ - HillFunction(0.0) - minSize;  // tanhFunction needs a plain operand
    for (int k = 0; k < NZ6 - 6; ++k) tz[k] = tanhFunction(total0, minSize, a, k, (NZ6 - 7));
    arr("tanh_column_y0", tz, NZ6 - 6);
    double txi[NZ6];
    for (int k = 0; k < NZ6 - 6; ++k) txi[k] = tanhFunction(LXi, minSize, a, k, (NZ6 - 7)) - minSize / 2.0;
    arr("xi_nodes", txi, NZ6 - 6);

    // GetParameter_6th (initializationTool.h:68-79) on the xi nodes and on a uniform grid
    const int nl = 9;
    static double lpos[nl], lw[nl * 7];
    static int lstart[nl];
    double* P[7];
    static double store[7][nl];
    for (int m = 0; m < 7; ++m) P[m] = store[m];
    for (int s = 0; s < nl; ++s) {
        lstart[s] = (s * 7) % (NZ6 - 6 - 7);
        lpos[s] = txi[lstart[s] + 3] + (s % 2 ? 0.37 : -0.61) * (txi[lstart[s] + 4] - txi[lstart[s] + 3]);
        GetParameter_6th(P, lpos[s], txi, s, lstart[s]);
        for (int m = 0; m < 7; ++m) lw[s * 7 + m] = P[m][s];
    }
    double lst[nl];
    for (int s = 0; s < nl; ++s) lst[s] = lstart[s];
    arr("lag_xi_start", lst, nl);
    arr("lag_xi_pos", lpos, nl);
    arr("lag_xi_w", lw, nl * 7);

    // BFL bisections (initializationTool.h:90-154) on the first two node layers of a
    // 128-plane streamwise grid, where the reference's BFL initialisation would call them.
    const double dy = LY / (double)NY;
    double hz[400], hyv[400], hd[400], fz[400], fyv[400], fd[400], gz[400], gyv[400], gd[400], kz[400], kyv[400], kd[400];
    int nh_ = 0, nf = 0, ng = 0, nk = 0;
    for (int j = 0; j < NY; ++j) {
        const double y = dy * j;
        const double H = HillFunction(y);
        const double total = LZ - H - minSize;
        for (int k = 0; k < 2; ++k) {
            const double z = tanhFunction(total, minSize, a, k, (NZ6 - 7)) + H;
            if (IsBFLBCNeeded(y - minSize, z) && nh_ < 400) {
                hz[nh_] = z; hyv[nh_] = y; hd[nh_++] = GetDeltaHorizontal(z, y - minSize, y, y);
            }
            if (IsBFLBCNeeded(y + minSize, z) && nf < 400) {
                fz[nf] = z; fyv[nf] = y; fd[nf++] = GetDeltaHorizontal(z, y + minSize, y, y);
            }
            if (IsBFLBCNeeded(y - minSize, z - minSize) && ng < 400) {
                gz[ng] = z; gyv[ng] = y; gd[ng++] = GetDelta45Degree(z, y, y, y - minSize);
            }
            if (IsBFLBCNeeded(y + minSize, z - minSize) && nk < 400) {
                kz[nk] = z; kyv[nk] = y; kd[nk++] = GetDelta45Degree(z, y, y, y + minSize);
            }
        }
    }
    arr("bfl3_z", hz, nh_); arr("bfl3_y", hyv, nh_); arr("bfl3_delta", hd, nh_);
    arr("bfl4_z", fz, nf); arr("bfl4_y", fyv, nf); arr("bfl4_delta", fd, nf);
    arr("bfl15_z", gz, ng); arr("bfl15_y", gyv, ng); arr("bfl15_delta", gd, ng);
    arr("bfl16_z", kz, nk); arr("bfl16_y", kyv, nk); arr("bfl16_delta", kd, nk);

    // MRT tables (MRT_Matrix.h:4-68)
    double s0, s1, s2, s3, s4, s5, s6, s7, s8, s9, s10, s11, s12, s13, s14, s15, s16, s17, s18;
    Relaxation;
    Matrix;
    Inverse_Matrix;
    double sv[19] = {s0, s1, s2, s3, s4, s5, s6, s7, s8, s9, s10, s11, s12, s13, s14, s15, s16, s17, s18};
    arr("relaxation", sv, 19);
    double mf[361], mif[361];
    for (int r = 0; r < 19; ++r)
        for (int c = 0; c < 19; ++c) { mf[r * 19 + c] = M[r][c]; mif[r * 19 + c] = Mi[r][c]; }
    arr("M", mf, 361);
    arr("Mi", mif, 361, true);
    std::printf("}\n");
    return 0;
}
