// Check program of the C++ layer include/agama_amd.hpp (built by tests/cpp/build_adapter_check.py,
// run by tests/test_cpp_adapter.py).
//   adapter_check errors            - error paths (no device needed): exceptions carry the message
//   adapter_check values COEF OUT   - potential, gradient and Hessian at 2000 points by eval() and
//                                     evalMany(), 64 orbits by integrateOrbits() and
//                                     integrateTrajectories(); checks their mutual consistency and
//                                     writes inputs and results to OUT for comparison with the oracle
#include <cstdio>
#include <cstring>
#include <cmath>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>
#include "agama_amd.hpp"

using namespace agama_amd;

static int errors()
{
    int bad = 0;
    try {
        MultipolePotential::create("type=Multipole, file=/nonexistent/coefs.ini");
        bad++;
    } catch(std::runtime_error& e) {
        if(!std::strstr(e.what(), "cannot open")) { std::printf("unexpected message: %s\n", e.what()); bad++; }
    }
    try {
        MultipolePotential::create("type=Multipole, density=Nonsense");
        bad++;
    } catch(std::runtime_error& e) {
        if(!std::strstr(e.what(), "unknown density")) { std::printf("unexpected message: %s\n", e.what()); bad++; }
    }
    std::printf(bad ? "errors FAILED\n" : "errors ok\n");
    return bad ? 1 : 0;
}

// deterministic points in a shell 0.01 < r < 50 (a multiplicative congruential sequence)
static double uniform(unsigned long long& s)
{
    s = s * 6364136223846793005ULL + 1442695040888963407ULL;
    return ((s >> 11) + 0.5) / 9007199254740992.0;
}

static int values(const char* coef, const char* out)
{
    std::shared_ptr<const MultipolePotential> pot = MultipolePotential::create(std::string("type=Multipole, file=") + coef);
    const size_t n = 2000;
    std::vector<PosCar> pts(n);
    unsigned long long s = 12345;
    for(size_t i = 0; i < n; i++) {
        const double r = std::exp(std::log(0.01) + uniform(s) * std::log(5000.0)), ct = 2 * uniform(s) - 1,
            ph = 2 * M_PI * uniform(s), st = std::sqrt(1 - ct * ct);
        pts[i] = PosCar{ r * st * std::cos(ph), r * st * std::sin(ph), r * ct };
    }
    std::vector<double> phi(n), phi2(n);
    std::vector<GradCar> g(n), g2(n);
    std::vector<HessCar> h(n);
    pot->evalMany(n, pts.data(), phi.data(), g.data(), h.data());
    pot->evalMany(n, pts.data(), phi2.data(), g2.data());
    int bad = 0;
    for(size_t i = 0; i < 50; i++) {   // single-point eval equals the batch of the same kernel
        double p1, p2;
        GradCar d1, d2;
        HessCar hh;
        pot->eval(pts[i], &p1, &d1);
        pot->eval(pts[i], &p2, &d2, &hh);
        if(p1 != phi2[i] || d1.dx != g2[i].dx || d1.dy != g2[i].dy || d1.dz != g2[i].dz) bad++;
        if(p2 != phi[i] || d2.dz != g[i].dz || hh.dxdz != h[i].dxdz || hh.dy2 != h[i].dy2) bad++;
        pot->eval(pts[i]);   // every output NULL
    }
    // orbits: near-circular starts in the x-y plane with a vertical kick, 3 circular periods
    const size_t m = 64;
    std::vector<PosVelCar> ic(m);
    std::vector<double> T(m);
    for(size_t i = 0; i < m; i++) {
        const double R = 0.1 * std::pow(200.0, i / (m - 1.0)), ph = 0.7 * i;
        double p;
        GradCar d;
        pot->eval(PosCar{ R * std::cos(ph), R * std::sin(ph), 0 }, &p, &d);
        const double vc = std::sqrt(R * std::hypot(d.dx, d.dy));
        ic[i] = PosVelCar{ R * std::cos(ph), R * std::sin(ph), 0.05 * R, -0.9 * vc * std::sin(ph), 0.9 * vc * std::cos(ph),
            0.2 * vc };
        T[i] = 3 * 2 * M_PI * R / vc;
    }
    OrbitBatchInfo info;
    std::vector<PosVelCar> fin = integrateOrbits(*pot, ic, T, 0.0, OrbitIntParams(1e-8), 0.0, &info);
    std::vector<Trajectory> traj = integrateTrajectories(*pot, ic, T, 5, 0.0, OrbitIntParams(1e-8));
    for(size_t i = 0; i < m; i++) {
        if(info.status[i] != 0 || traj[i].size() != 5) { bad++; continue; }
        const PosVelCar& a = traj[i].back().first;
        if(a.x != fin[i].x || a.vz != fin[i].vz || traj[i].back().second != T[i]) bad++;
        if(traj[i].front().first.x != ic[i].x || traj[i].front().second != 0) bad++;
    }
    try {   // mismatched arrays are refused before anything runs
        integrateOrbits(*pot, ic, std::vector<double>(3, 1.0));
        bad++;
    } catch(std::invalid_argument&) {}
    std::ofstream f(out, std::ios::binary);
    auto put = [&](const void* p, size_t bytes) { f.write(static_cast<const char*>(p), (std::streamsize)bytes); };
    put(pts.data(), n * 24); put(phi.data(), n * 8); put(g.data(), n * 24); put(h.data(), n * 48);
    put(ic.data(), m * 48); put(T.data(), m * 8); put(fin.data(), m * 48); put(info.nsteps.data(), m * 8);
    std::printf(bad ? "values FAILED (%d)\n" : "values ok\n", bad);
    return bad ? 1 : 0;
}

int main(int argc, char** argv)
{
    try {
        if(argc >= 2 && !std::strcmp(argv[1], "errors")) return errors();
        if(argc >= 4 && !std::strcmp(argv[1], "values")) return values(argv[2], argv[3]);
    } catch(std::exception& e) {
        std::printf("exception: %s\n", e.what());
        return 2;
    }
    std::printf("usage: adapter_check errors | values COEF OUT\n");
    return 3;
}
