/*  AddressSanitizer / UndefinedBehaviorSanitizer run of the CPU code - TEST INFRASTRUCTURE.

    Built by `make -C oracle sanitize` from the oracle (agama_oracle.cpp) and the library's host
    construction code (agama_amd/csrc/host_multipole.cpp, host_lyapunov.cpp) with
    -fsanitize=address,undefined -fno-sanitize-recover=all, then run by tests/test_sanitize.py:
    any invalid access, leak, overflow or other undefined behaviour ends it with a report and a
    non-zero status.  It drives every host routine of the library (coefficient files, Poisson
    construction, 2-D / 1-D tables, per-cell polynomials, bucket and Legendre tables, the
    Lyapunov fit) and the oracle's evaluation and orbit routines on the golden fixtures, and
    checks a few results against each other on the way (the host tables' grids equal the
    oracle's; a coefficient file read back equals what was written).

    Usage: sanitize_check <tests/golden directory> <scratch file path>
           sanitize_check --planted-overflow   (reads one element past a heap array: shows that
                                                the sanitizers are live; must fail)
*/
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>
#include "agama_oracle.h"
#include "../agama_amd/csrc/host_multipole.h"
#include "../agama_amd/csrc/host_lyapunov.h"

static int g_fail = 0;
static void check(bool ok, const char* what)
{
    if(!ok) { std::fprintf(stderr, "FAILED: %s\n", what); g_fail = 1; }
}

// points inside the grid, beyond both ends, on the z axis and at the origin
static std::vector<double> probePoints()
{
    std::vector<double> p;
    const double rr[] = { 1e-5, 3e-3, 0.1, 1.0, 7.0, 90.0, 2e3, 1e6 };
    for(double r : rr)
        for(int k = 0; k < 5; k++) {
            const double th = 0.3 + 0.6 * k, ph = 0.7 * k;
            p.push_back(r * std::sin(th) * std::cos(ph));
            p.push_back(r * std::sin(th) * std::sin(ph));
            p.push_back(r * std::cos(th));
        }
    p.insert(p.end(), { 0, 0, 0.5, 0, 0, -2, 0, 0, 0 });
    return p;
}

static void hostTables(const std::string& file)
{
    const agx::Coefs c = agx::readCoefFile(file);
    const agx::HostTables T = agx::buildTables(c);
    if(!T.interp1d) {
        const std::vector<double> cells = agx::buildCellPolys(T);
        const std::vector<int> lut = agx::buildTauLookup(T);
        check(cells.size() == (size_t)(T.nr - 1) * (T.nt - 1) * T.ns * AGX_CELL_COEFS, "cell table size");
        for(int b : lut) check(b >= 0 && b < T.nt - 1, "bucket table entry in range");
        // the oracle's grids are the same numbers
        orc_pot* p = orc_multipole_from_file(file.c_str());
        check(p != nullptr, "oracle potential");
        if(p) {
            const long nr = orc_multipole_tables(p, 0, 0, nullptr), nt = orc_multipole_tables(p, 1, 0, nullptr);
            std::vector<double> gr(nr > 0 ? nr : 0), gt(nt > 0 ? nt : 0);
            orc_multipole_tables(p, 0, 0, gr.data());
            orc_multipole_tables(p, 1, 0, gt.data());
            check(gr == T.gridR && gt == T.gridT, "host grids equal the oracle's");
            orc_free(p);
        }
    }
    const std::vector<double> leg = agx::buildLegendreTable(T.lmax);
    check(!leg.empty(), "Legendre table");
    const agx::HostTables T2 = agx::buildTables(agx::truncateCoefs(c, 2), true);
    check(T2.nr == T.nr, "truncated expansion's tables");
}

static void coefFileRoundTrip(const std::string& file, const std::string& scratch)
{
    const agx::Coefs c = agx::readCoefFile(file);
    agx::writeCoefFile(scratch, c);
    const agx::Coefs d = agx::readCoefFile(scratch);
    check(d.radii == c.radii && d.Phi == c.Phi && d.dPhi == c.dPhi, "coefficient file round trip");
}

static void construction(const std::string& scratch)
{
    const char* models[] = {
        "density=Dehnen, mass=1, scaleRadius=1, gamma=1, axisRatioY=0.8, axisRatioZ=0.6",
        "density=Ferrers, mass=0.1, scaleRadius=2, axisRatioY=0.5, axisRatioZ=0.3",
        "density=Disk+Dehnen, 0.mass=0.05, 0.scaleRadius=3, 0.scaleHeight=0.3, 1.mass=1, 1.scaleRadius=2",
        "density=Plummer, mass=1, scaleRadius=1" };
    for(const char* m : models) {
        const agx::Params P = agx::Params::parse(m);
        auto dens = agx::makeDensity(P);
        const int lmax = 6;
        const agx::Coefs c = agx::multipoleFromDensity(*dens, lmax, lmax, dens->symmetry(), 20, 0.01, 100.);
        check(c.radii.size() == 20, "Poisson construction grid");
        // the staged form used by the device solver, run on the host
        const agx::PoissonSetup S = agx::poissonSetup(*dens, lmax, lmax, dens->symmetry(), 20, 0.01, 100.);
        std::vector<double> Phi, dPhi;
        agx::poissonStagesHost(S, *dens, Phi, dPhi);
        const agx::Coefs c2 = agx::poissonFinish(S, Phi, dPhi);
        check(c2.radii == c.radii, "staged Poisson construction grid");
        agx::writeCoefFile(scratch, c);
        const agx::HostTables T = agx::buildTables(agx::readCoefFile(scratch));
        check(T.nr == 20, "tables of the constructed expansion");
    }
}

static void oracleRuns(const std::string& file)
{
    orc_pot* p = orc_multipole_from_file(file.c_str());
    check(p != nullptr, "oracle potential");
    if(!p) return;
    const std::vector<double> pts = probePoints();
    const long n = (long)pts.size() / 3;
    std::vector<double> phi(n), grad(3 * n), hess(6 * n);
    orc_eval_many(p, n, pts.data(), 0.0, phi.data(), grad.data(), 2);
    orc_eval_hess_many(p, n, pts.data(), 0.0, phi.data(), grad.data(), hess.data(), 2);
    // a few orbits: trajectories, DOP853 and DPRKN8, the rotating frame
    const int no = 6;
    std::vector<double> ic(6 * no), T(no);
    for(int i = 0; i < no; i++) {
        const double r = 0.2 * (i + 1);
        ic[6 * i] = r; ic[6 * i + 1] = 0.1 * r; ic[6 * i + 2] = 0.05 * i;
        ic[6 * i + 3] = 0.02; ic[6 * i + 4] = 0.5 / std::sqrt(r); ic[6 * i + 5] = 0.1;
        T[i] = 5.0 + i;
    }
    std::vector<int> sz(no, 4);
    std::vector<long> off(no);
    for(int i = 0; i < no; i++) off[i] = 4L * i;
    std::vector<double> tout(4 * no), traj(24 * no), fin(6 * no), maxde(no);
    std::vector<long> ns(no), na(no), ne(no);
    std::vector<int> st(no);
    for(int method = 0; method < 2; method++)
        for(double Omega : { 0.0, 0.3 }) {
            const int rc = orc_orbit_many_method(p, no, ic.data(), nullptr, T.data(), sz.data(), off.data(), 1e-8,
                Omega, 100000, tout.data(), traj.data(), fin.data(), ns.data(), na.data(), ne.data(),
                maxde.data(), st.data(), 2, method);
            check(rc == 0, "oracle orbits");
            for(int i = 0; i < no; i++) check(st[i] == 0 && maxde[i] < 1e-4, "orbit status and energy");
        }
    orc_free(p);
}

static void lyapunovFit()
{
    // exponential growth after a linear phase: the fit finds a positive exponent
    std::vector<double> t, lnw;
    for(int k = 0; k < 40; k++) {
        const double tk = 10.0 * (k + 1);
        t.push_back(tk);
        lnw.push_back(tk < 150 ? std::log(tk) : std::log(150.0) + 0.05 * (tk - 150));
    }
    double out[2];
    agx::lyapunovFromBins(t.data(), lnw.data(), (int)t.size(), 400.0, 10.0, out);
    check(out[0] > 0, "Lyapunov exponent of exponential growth");
    agx::lyapunovFromBins(t.data(), lnw.data(), 3, 30.0, 10.0, out);
    check(std::isnan(out[0]), "too few bins");
}

int main(int argc, char** argv)
{
    if(argc == 2 && std::string(argv[1]) == "--planted-overflow") {
        std::vector<double> v(7, 1.0);
        volatile const double* q = v.data();
        std::printf("%g\n", q[v.size()]);   // one past the end
        return 0;
    }
    if(argc < 3) { std::fprintf(stderr, "usage: %s <golden dir> <scratch file>\n", argv[0]); return 2; }
    const std::string dir = argv[1], scratch = argv[2];
    const char* files[] = { "dehnen_triax_lmax8.coef", "dehnen_triax_lmax2.coef", "disk_bulge_halo_lmax12.coef",
        "ferrers_bar_lmax16.coef", "plummer_lmax4.coef" };
    for(const char* f : files) {
        const std::string path = dir + "/" + f;
        hostTables(path);
        coefFileRoundTrip(path, scratch);
        oracleRuns(path);
        std::printf("%s: done\n", f);
    }
    construction(scratch);
    std::printf("construction: done\n");
    lyapunovFit();
    std::printf("lyapunov: done\n");
    std::remove(scratch.c_str());
    return g_fail;
}
