// C ABI of libagama_amd (declared in include/agama_amd.h): potential handles, table upload,
// batched evaluation and orbit integration launchers, the NEMO plugin entry point and the
// drop-in aliases of the reference's plain C interface.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>
#include "../../include/agama_amd.h"
#include "agx_kernels.h"
#include "agx_surrogate.h"
#include "host_multipole.h"
#include "host_lyapunov.h"
#include "agx_target.h"

hipError_t agxLaunchEval(const AgxMultipoleDev& P, const AgxMultipoleDev* dP, int precision, int fetch,
    long n, const double* pos, double* phi, double* grad, hipStream_t stream);
hipError_t agxLaunchOrbit(const AgxMultipoleDev& P, const AgxMultipoleDev* dP, int precision, int fetch,
    int waves, const AgxOrbitArgs& A, int numBlocks, int blockSize, hipStream_t stream, AgxOrbitArgs* dArgs);
int agxOrbitKernelBlocksPerCU(int precision, int fetch, int waves, bool nfw, bool targets = false);
hipError_t agxLaunchOrbitRKN(const AgxMultipoleDev& P, const AgxMultipoleDev* dP, int precision, int fetch,
    const AgxOrbitArgs& A, int numBlocks, hipStream_t stream);
int agxOrbitKernelRKNBlocksPerCU(int precision, bool nfw);
bool agxOrbitRKNCellsFit(const AgxMultipoleDev& P);
int agxOrbitVarKernelBlocksPerCU(bool nfw);
size_t agxOrbitVarStateDoubles(int numBlocks);
hipError_t agxLaunchOrbitVar(const AgxMultipoleDev& P, const AgxMultipoleDev* dP, const AgxOrbitArgs& A,
    int numBlocks, double* varState, double* dev, const double* period, const double* maxTime, double* bins,
    int* nbins, double* lastTime, int maxBins, hipStream_t stream);
hipError_t agxLaunchTcirc(const AgxMultipoleDev* dP, long n, const double* ic, const double* tstart,
    double Omega, double* period, hipStream_t stream);
hipError_t agxLaunchFillU32(unsigned int* p, long n, unsigned int v, hipStream_t stream);
hipError_t agxLaunchEvalHess(const AgxMultipoleDev* dP, long n, const double* pos, double* phi,
    double* grad, double* hess, hipStream_t stream);
hipError_t agxLaunchBuildSplines(const AgxSeries* dSeries, long ns, double Omega, double* nodes,
    double* pivots, int* count, int* reversed, hipStream_t stream);
hipError_t agxLaunchEvalSplines(const AgxSeries* dSeries, const int* count, const double* nodes, double Omega,
    long m, const long* which, const double* t, int mode, double* out, hipStream_t stream);
hipError_t agxLaunchGatherRows(long nseg, const long* srcRow, const long* dstRow, const long* rows, int width,
    const double* src, double* dst, hipStream_t stream);
hipError_t agxPoissonDevice(const agx::PoissonSetup& S, const agx::Density& dens, std::vector<double>& Phi,
    std::vector<double>& dPhi);
int agxOrbitKernelBlockSize();
size_t agxScheduleWorkspace(long n);
double* agxScheduleProbeTimes(void* work, long n);
hipError_t agxLaunchScheduleFromProbe(long n, const double* tstart, const double* ttotal, int probeSteps,
    int* order, void* work, int** finalOrder, hipStream_t stream);
hipError_t agxLaunchSchedule(const AgxMultipoleDev& P, const AgxMultipoleDev* dP, long n,
    const double* ic, const double* ttotal, int* order, void* work, hipStream_t stream);
hipError_t agxLaunchDealGroups(int* order, long groupOrbits, hipStream_t stream);

struct agx_potential {
    agx::Coefs coefs;
    agx::HostTables T;
    AgxMultipoleDev dev;               // header with device pointers (host copy)
    AgxMultipoleDev* ddev = nullptr;   // the same header in device memory
    void* dmem = nullptr;
    unsigned int* dcounter = nullptr;
    AgxOrbitArgs* dargs = nullptr;     // device copy of the orbit kernel's arguments (per launch)
    int device = 0;
    double nfwMass = 0, nfwScale = 1;  // optional analytic NFW component
    // staging of the host-buffer evaluation entry points (agx_eval_many, agx_eval_hess_many):
    // the handle's own stream and a device buffer kept between calls (grown to the largest
    // call), so that a call neither allocates nor waits for other streams; calls on one handle
    // take the mutex in turn
    mutable std::mutex evalMutex;
    mutable hipStream_t evalStream = nullptr;
    mutable double* evalBuf = nullptr;
    mutable size_t evalCap = 0;
    ~agx_potential()
    {
        if(evalBuf) (void)hipFree(evalBuf);
        if(evalStream) (void)hipStreamDestroy(evalStream);
        if(dmem) (void)hipFree(dmem);
    }
};

namespace {

thread_local std::string g_err;
int g_orbitBlocks = 0;
int g_fetchEval = AGX_FETCH_CELL, g_fetchOrbit = AGX_FETCH_CELL;   // set by agx_set_fetch
int g_orbitWaves = 1;                                                // set by agx_set_orbit_waves
// group mode of the DOP853 kernel (agx_set_orbit_groups): one wavefront in g_groupStride
// integrates the longest orbits with 8 lanes each, for batches of at least g_groupMinBatch;
// stride -1: chosen from the orbits per lane of the launch
int g_groupStride = -1;
long g_groupOrbits = -1, g_groupMinBatch = 65536;

int fail(int code, const std::string& msg) { g_err = msg; return code; }

// the surrogate orbit batch's pilot (agx_set_surrogate_pilot) and its last outcome
// work-order probe (agx_set_orbit_probe): steps per orbit (0: off) and the smallest batch probed
int g_probeSteps = 64;
long g_probeMinBatch = 65536;

int g_surrogatePilot = 1;
thread_local long g_lastPilot[2] = { 0, 0 };   // pilot orbits, of which fell back

// where Multipole coefficients of a density are computed (agx_set_construction); -1: not set,
// the device for agx_potential_create and the host for agx_build_coef_file
int g_construction = -1;

struct HipError : std::runtime_error {
    explicit HipError(const std::string& s) : std::runtime_error(s) {}
};
#define HIPCHECK(expr) do { hipError_t e_ = (expr); if(e_ != hipSuccess) \
    throw HipError(std::string(#expr) + ": " + hipGetErrorString(e_)); } while(0)

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// upload the host tables into one device allocation and fill the kernel-argument header
void upload(agx_potential& p)
{
    const agx::HostTables& T = p.T;
    const size_t nc = p.coefs.Phi.size();
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += align256(bytes); return o; };
    const size_t oR = take(T.gridR.size() * 8), oT = take(T.gridT.size() * 8),
        oN = take(T.nodes.size() * 8), oF = take(T.nodes.size() * 4),
        oIS = take(nc * 8), oIU = take(nc * 8), oIW = take(nc * 8),
        oOS = take(nc * 8), oOU = take(nc * 8), oOW = take(nc * 8), oC = take(64),
        oN1 = take(T.nodes1d.size() * 8), oA = take(sizeof(AgxOrbitArgs)), oH = take(sizeof(AgxMultipoleDev));
    const std::vector<double> cells = agx::buildCellPolys(T);
    if(cells.size() * 8 >= (1ull << 31)) throw std::runtime_error("Multipole tables too large for 32-bit offsets");
    const size_t oCell = take(cells.size() * 8);
    const std::vector<int> tauLut = agx::buildTauLookup(T);
    const size_t oLut = take(tauLut.size() * 4);
    const int legLmax = std::max(T.inner.lmax, T.outer.lmax);
    const std::vector<double> legTab = agx::buildLegendreTable(legLmax);
    const size_t oLeg = take(legTab.size() * 8);
    HIPCHECK(hipGetDevice(&p.device));
    HIPCHECK(hipMalloc(&p.dmem, off));
    char* base = static_cast<char*>(p.dmem);
    std::vector<float> nf(T.nodes.begin(), T.nodes.end());
    auto put = [&](size_t o, const void* src, size_t bytes) {
        if(bytes) HIPCHECK(hipMemcpy(base + o, src, bytes, hipMemcpyHostToDevice));
    };
    put(oR, T.gridR.data(), T.gridR.size() * 8);
    put(oT, T.gridT.data(), T.gridT.size() * 8);
    put(oN, T.nodes.data(), T.nodes.size() * 8);
    put(oF, nf.data(), nf.size() * 4);
    put(oIS, T.inner.S.data(), nc * 8); put(oIU, T.inner.U.data(), nc * 8); put(oIW, T.inner.W.data(), nc * 8);
    put(oOS, T.outer.S.data(), nc * 8); put(oOU, T.outer.U.data(), nc * 8); put(oOW, T.outer.W.data(), nc * 8);
    put(oN1, T.nodes1d.data(), T.nodes1d.size() * 8);
    put(oCell, cells.data(), cells.size() * 8);
    put(oLut, tauLut.data(), tauLut.size() * 4);
    put(oLeg, legTab.data(), legTab.size() * 8);
    HIPCHECK(hipMemset(base + oC, 0, 64));
    AgxMultipoleDev& D = p.dev;
    memset(&D, 0, sizeof(D));
    D.nr = T.nr; D.nt = T.nt; D.ns = T.ns; D.mmax = T.mmax; D.mmin = T.mmin; D.sym = T.sym;
    D.logScaling = T.logScaling; D.invPhi0 = T.invPhi0;
    D.rsqInner = T.rsqInner; D.rsqOuter = T.rsqOuter;
    for(int s = 0; s < T.ns; s++) D.slotM[s] = T.slotM[s];
    D.zeroSlot = T.zeroSlot;
    D.nfwMass = p.nfwMass;
    D.nfwScale = p.nfwScale;
    D.gridR = reinterpret_cast<const double*>(base + oR);
    D.gridT = reinterpret_cast<const double*>(base + oT);
    D.nodes = reinterpret_cast<const double*>(base + oN);
    D.nodesF = reinterpret_cast<const float*>(base + oF);
    D.cells = cells.empty() ? nullptr : reinterpret_cast<const double*>(base + oCell);
    D.cellBytes = (unsigned int)(cells.size() * 8);
    D.lnrScale = T.nr > 1 ? (T.nr - 1) / (T.gridR.back() - T.gridR.front()) : 0;
    D.nTauLut = (int)tauLut.size();
    D.tauLut = tauLut.empty() ? nullptr : reinterpret_cast<const int*>(base + oLut);
    D.legLmax = legLmax;
    D.legTab = reinterpret_cast<const double*>(base + oLeg);
    D.legStart = D.legTab + (size_t)(legLmax + 1) * (legLmax + 1) * AGX_LEG_STRIDE;
    auto fillPL = [&](AgxPowerLawDev& L, const agx::HostPowerLaw& H, size_t s, size_t u, size_t w) {
        L.r0sq = H.r0sq; L.Q = H.Q; L.inner = H.inner; L.lmax = H.lmax; L.mmax = H.mmax;
        L.mmin = H.mmin; L.step = H.step;
        for(int i = 0; i < AGX_MAX_SLOTS; i++) L.lmin[i] = i < (int)H.lmin.size() ? H.lmin[i] : 1 << 20;
        L.uniformS = 1;   // over the terms of the index set
        for(int m = H.mmin; m <= H.mmax; m++) {
            const int lm = H.lmin[m + H.mmax];
            for(int l = lm; l <= H.lmax; l += H.step) {
                const int c = l * (l + 1) + m;
                if(c < (int)H.S.size() && H.S[c] != (H.inner ? l : -l - 1) && H.S[c] != H.S[0]) L.uniformS = 0;
            }
        }
        L.S = reinterpret_cast<const double*>(base + s);
        L.U = reinterpret_cast<const double*>(base + u);
        L.W = reinterpret_cast<const double*>(base + w);
    };
    fillPL(D.inner, T.inner, oIS, oIU, oIW);
    fillPL(D.outer, T.outer, oOS, oOU, oOW);
    D.interp1d = T.interp1d;
    D.nc1d = T.nc1d;
    D.nodes1d = reinterpret_cast<const double*>(base + oN1);
    D.ind1d.lmax = T.lmax1d; D.ind1d.mmax = T.mmax1d; D.ind1d.mmin = T.mmin1d; D.ind1d.step = T.step1d;
    for(int i = 0; i < AGX_MAX_SLOTS; i++) D.ind1d.lmin[i] = i < (int)T.lmin1d.size() ? T.lmin1d[i] : 1 << 20;
    p.dcounter = reinterpret_cast<unsigned int*>(base + oC);
    p.dargs = reinterpret_cast<AgxOrbitArgs*>(base + oA);
    p.ddev = reinterpret_cast<AgxMultipoleDev*>(base + oH);
    put(oH, &D, sizeof(D));
}

// the Poisson solver for a density: the bulk on the device (default), or on the host in the
// reference's order or in the device's order (agx_set_construction)
agx::Coefs coefsFromDensity(const agx::Density& dens, int lmax, int mmax, int sym, int nr, double rmin, double rmax,
    int mode)
{
    if(mode == AGX_CONSTRUCT_HOST)
        return agx::multipoleFromDensity(dens, lmax, mmax, sym, nr, rmin, rmax);
    // the grids and the transform matrix depend on the density only through its symmetry:
    // kept from the previous construction with the same parameters (rebuilding a potential in a
    // loop then costs the device part and the tables only)
    static std::mutex setupLock;
    static std::shared_ptr<const agx::PoissonSetup> lastSetup;
    static std::vector<double> lastKey;
    const std::vector<double> key = { (double)dens.symmetry(), (double)lmax, (double)mmax, (double)sym, (double)nr,
        rmin, rmax };
    std::shared_ptr<const agx::PoissonSetup> Sp;
    {
        std::lock_guard<std::mutex> g(setupLock);
        if(!lastSetup || lastKey != key) {
            lastSetup = std::make_shared<const agx::PoissonSetup>(agx::poissonSetup(dens, lmax, mmax, sym, nr, rmin, rmax));
            lastKey = key;
        }
        Sp = lastSetup;
    }
    const agx::PoissonSetup& S = *Sp;
    std::vector<double> Phi, dPhi;
    if(mode == AGX_CONSTRUCT_HOST_STAGED)
        agx::poissonStagesHost(S, dens, Phi, dPhi);
    else {
        const hipError_t e = agxPoissonDevice(S, dens, Phi, dPhi);
        if(e != hipSuccess)
            throw std::runtime_error(std::string("Multipole construction on the device: ") + hipGetErrorString(e));
    }
    return agx::poissonFinish(S, Phi, dPhi);
}

agx_potential* makeFromCoefs(agx::Coefs&& c, bool allowLowOrder = false, double nfwMass = 0,
    double nfwScale = 1)
{
    std::unique_ptr<agx_potential> p(new agx_potential);
    p->nfwMass = nfwMass;
    p->nfwScale = nfwScale;
    p->coefs = std::move(c);
    p->T = agx::buildTables(p->coefs, allowLowOrder);
    upload(*p);
    return p.release();
}

// "type=Multipole, file=..." or "type=Multipole, density=..., lmax=..., gridSizeR=..., ...",
// optionally followed by "; type=NFW, mass=..., scaleRadius=..." (a two-component Composite,
// src/potential_composite.cpp; components separated by ';')
agx_potential* createFromParams(const std::string& text)
{
    std::vector<std::string> parts;
    size_t b = 0;
    for(size_t e; (e = text.find(';', b)) != std::string::npos; b = e + 1) parts.push_back(text.substr(b, e - b));
    parts.push_back(text.substr(b));
    const agx::Params* mp = nullptr;
    std::vector<agx::Params> comps;
    for(const std::string& t : parts) comps.push_back(agx::Params::parse(t));
    double nfwMass = 0, nfwScale = 1;
    int nfwCount = 0;
    for(const agx::Params& P : comps) {
        std::string tl = P.str("type", "Multipole");
        for(auto& ch : tl) ch = (char)tolower(ch);
        if(tl == "multipole") {
            if(mp) throw std::invalid_argument("a composite may hold only one Multipole component");
            mp = &P;
        } else if(tl == "nfw") {
            P.checkOnly({"type", "mass", "scaleradius"});
            nfwMass = P.num("mass", 1);
            nfwScale = P.num("scaleradius", 1);
            if(!(nfwScale > 0) || !(nfwMass > -INFINITY && nfwMass < INFINITY))
                throw std::invalid_argument("NFW: invalid mass or scaleRadius");
            if(++nfwCount > 1) throw std::invalid_argument("a composite may hold only one NFW component");
        } else
            throw std::invalid_argument("only type=Multipole (plus an optional type=NFW component) is "
                "provided by this library (got " + P.str("type") + ")");
    }
    if(!mp) throw std::invalid_argument("a Multipole component is required");
    const agx::Params& P = *mp;
    if(P.has("file")) {
        P.checkOnly({"type", "file"});
        return makeFromCoefs(agx::readCoefFile(P.str("file")), false, nfwMass, nfwScale);
    }
    if(!P.has("density"))
        throw std::invalid_argument("Multipole needs either file=... or density=...");
    P.checkOnly({"type", "density", "lmax", "mmax", "gridsizer", "rmin", "rmax", "symmetry",
        "mass", "scaleradius", "gamma", "axisratioy", "axisratioz", "scaleheight", "surfacedensity",
        "0.*", "1.*", "2.*", "3.*", "4.*", "5.*"});
    auto dens = agx::makeDensity(P);
    const int lmax = (int)P.num("lmax", 6);
    const int mmax = (int)P.num("mmax", lmax);
    const int nr = (int)P.num("gridsizer", 25);
    const int sym = agx::symmetryFromName(P.str("symmetry"));
    return makeFromCoefs(coefsFromDensity(*dens, lmax, mmax, sym, nr, P.num("rmin", 0), P.num("rmax", 0),
        g_construction < 0 ? AGX_CONSTRUCT_DEVICE : g_construction), false, nfwMass, nfwScale);
}

int orbitGeometry(long n, int precision, bool nfw, int& blocks, int& bs, bool rkn = false, bool targets = false)
{
    bs = agxOrbitKernelBlockSize();   // the DOP853 and DPRKN8 kernels share the block size
    if(g_orbitBlocks > 0) blocks = g_orbitBlocks;
    else {
        int dev = 0, cus = 0;
        HIPCHECK(hipGetDevice(&dev));
        HIPCHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        const int perCU = std::max(1, rkn ? agxOrbitKernelRKNBlocksPerCU(precision, nfw) :
            agxOrbitKernelBlocksPerCU(precision, g_fetchOrbit, g_orbitWaves, nfw, targets));
        blocks = cus * perCU;
    }
    const long need = (n + bs - 1) / bs;
    if(need < blocks) blocks = (int)std::max<long>(need, 1);
    return 0;
}

// the persistent orbit kernel of the method named in flags (DOP853 by default)
// dArgs: where the DOP853 kernel's arguments are placed (default: the potential's slot)
hipError_t launchOrbitMethod(const agx_potential* pot, int precision, int flags, const AgxOrbitArgs& A,
    int blocks, int bs, hipStream_t st, AgxOrbitArgs* dArgs = nullptr)
{
    if(flags & AGX_METHOD_DPRKN8) return agxLaunchOrbitRKN(pot->dev, pot->ddev, precision, g_fetchOrbit, A, blocks, st);
    return agxLaunchOrbit(pot->dev, pot->ddev, precision, g_fetchOrbit, g_orbitWaves, A, blocks, bs, st,
        dArgs ? dArgs : pot->dargs);
}

template<typename T>
struct DevBuf {
    T* p = nullptr;
    explicit DevBuf(size_t n) { if(n) HIPCHECK(hipMalloc(&p, n * sizeof(T))); }
    ~DevBuf() { if(p) (void)hipFree(p); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
};

}  // namespace

extern "C" {

const char* agx_last_error(void) { return g_err.c_str(); }

agx_potential* agx_potential_create(const char* params)
{
    g_err.clear();
    if(!params) { fail(AGX_EINVAL, "params is NULL"); return nullptr; }
    try { return createFromParams(params); }
    catch(std::exception& e) { fail(AGX_EINVAL, e.what()); return nullptr; }
}

agx_potential* agx_potential_from_coefs(int nr, int lmax, const double* radii,
    const double* phi, const double* dphi)
{
    g_err.clear();
    try {
        if(nr < 2 || lmax < 0 || !radii || !phi || !dphi) throw std::invalid_argument("invalid coefficient arrays");
        agx::Coefs c;
        c.radii.assign(radii, radii + nr);
        const int nc = (lmax + 1) * (lmax + 1);
        c.Phi.resize(nc); c.dPhi.resize(nc);
        for(int i = 0; i < nc; i++) {
            c.Phi[i].assign(phi + (size_t)i * nr, phi + (size_t)(i + 1) * nr);
            c.dPhi[i].assign(dphi + (size_t)i * nr, dphi + (size_t)(i + 1) * nr);
        }
        return makeFromCoefs(std::move(c));
    } catch(std::exception& e) { fail(AGX_EINVAL, e.what()); return nullptr; }
}

void agx_potential_destroy(agx_potential* pot) { delete pot; }

int agx_build_coef_file(const char* params, const char* path)
{
    g_err.clear();
    if(!params || !path) return fail(AGX_EINVAL, "NULL argument");
    try {
        agx::Params P = agx::Params::parse(params);
        if(!P.has("density")) throw std::invalid_argument("density=... is required");
        P.checkOnly({"type", "density", "lmax", "mmax", "gridsizer", "rmin", "rmax", "symmetry",
            "mass", "scaleradius", "gamma", "axisratioy", "axisratioz", "scaleheight", "surfacedensity",
            "0.*", "1.*", "2.*", "3.*", "4.*", "5.*"});
        auto dens = agx::makeDensity(P);
        const int lmax = (int)P.num("lmax", 6);
        agx::Coefs c = coefsFromDensity(*dens, lmax, (int)P.num("mmax", lmax),
            agx::symmetryFromName(P.str("symmetry")), (int)P.num("gridsizer", 25),
            P.num("rmin", 0), P.num("rmax", 0), g_construction < 0 ? AGX_CONSTRUCT_HOST : g_construction);
        agx::writeCoefFile(path, c);
        return AGX_OK;
    } catch(std::exception& e) { return fail(AGX_EINVAL, e.what()); }
}

int agx_set_construction(int mode)
{
    if(mode != AGX_CONSTRUCT_DEVICE && mode != AGX_CONSTRUCT_HOST && mode != AGX_CONSTRUCT_HOST_STAGED &&
        mode != AGX_CONSTRUCT_DEFAULT)
        return fail(AGX_EINVAL, "unknown construction mode");
    g_construction = mode;
    return AGX_OK;
}

long agx_debug_tables(const char* coef_path, int which, int m, double* out)
{
    g_err.clear();
    try {
        agx::HostTables T = agx::buildTables(agx::readCoefFile(coef_path));
        std::vector<double> v;
        if(which == 0) v = T.gridR;
        else if(which == 1) v = T.gridT;
        else if(which == 2) {
            auto it = std::find(T.slotM.begin(), T.slotM.end(), m);
            if(it == T.slotM.end()) return 0;
            const int s = (int)(it - T.slotM.begin());
            const size_t n2 = (size_t)T.nr * T.nt;
            v.resize(9 * n2);
            for(int q = 0; q < 9; q++)
                for(size_t i = 0; i < n2; i++)
                    v[q * n2 + i] = T.nodes[(i * T.ns + s) * AGX_NODE_STRIDE + q];
        } else if(which == 3)
            v = { (double)T.logScaling, T.invPhi0, (double)T.lmax, (double)T.mmax, (double)T.mmin, (double)T.step };
        else if(which == 5) {   // per-cell polynomial form of harmonic m: 36 coefficients per cell
            auto it = std::find(T.slotM.begin(), T.slotM.end(), m);
            if(T.interp1d || it == T.slotM.end()) return 0;
            const int s = (int)(it - T.slotM.begin());
            const std::vector<double> cells = agx::buildCellPolys(T);
            const size_t nc = (size_t)(T.nr - 1) * (T.nt - 1);
            v.resize(nc * AGX_CELL_COEFS);
            for(size_t c = 0; c < nc; c++)
                std::copy_n(&cells[(c * T.ns + s) * AGX_CELL_COEFS], AGX_CELL_COEFS, &v[c * AGX_CELL_COEFS]);
        }
        else if(which == 4) {   // 1-d variant: term c = m of the call: f, f_x, f_xx at every node
            if(!T.interp1d || m < 0 || m >= T.nc1d || !T.term1d[m]) return 0;
            v.resize(3 * (size_t)T.nr);
            for(int q = 0; q < 3; q++)
                for(int k = 0; k < T.nr; k++) v[q * T.nr + k] = T.nodes1d[((size_t)k * T.nc1d + m) * 4 + q];
        } else return -1;
        if(out) std::copy(v.begin(), v.end(), out);
        return (long)v.size();
    } catch(std::exception& e) { fail(AGX_EINVAL, e.what()); return -1; }
}

int agx_potential_write(const agx_potential* pot, const char* path)
{
    g_err.clear();
    if(!pot || !path) return fail(AGX_EINVAL, "NULL argument");
    try { agx::writeCoefFile(path, pot->coefs); return AGX_OK; }
    catch(std::exception& e) { return fail(AGX_EINVAL, e.what()); }
}

int agx_potential_coefs(const agx_potential* pot, int* nr, int* lmax, double* radii,
    double* phi, double* dphi)
{
    g_err.clear();
    if(!pot) return fail(AGX_EINVAL, "NULL potential");
    const int n = (int)pot->coefs.radii.size(), L = pot->coefs.lmax();
    if(nr) *nr = n;
    if(lmax) *lmax = L;
    if(radii) std::copy(pot->coefs.radii.begin(), pot->coefs.radii.end(), radii);
    for(size_t c = 0; c < pot->coefs.Phi.size(); c++) {
        if(phi) std::copy(pot->coefs.Phi[c].begin(), pot->coefs.Phi[c].end(), phi + c * n);
        if(dphi) std::copy(pot->coefs.dPhi[c].begin(), pot->coefs.dPhi[c].end(), dphi + c * n);
    }
    return AGX_OK;
}

int agx_eval_many_device(const agx_potential* pot, long n, const double* d_pos, double t,
    double* d_phi, double* d_grad, int flags, void* stream)
{
    (void)t;   // the Multipole potential is static
    if(!pot) return fail(AGX_EINVAL, "NULL potential");
    if(n < 0 || (n > 0 && !d_pos)) return fail(AGX_EINVAL, "invalid point array");
    if(flags != AGX_PRECISION_FP64 && flags != AGX_PRECISION_FP32) return fail(AGX_EINVAL, "invalid flags");
    hipError_t e = agxLaunchEval(pot->dev, pot->ddev, flags, g_fetchEval, n, d_pos, d_phi, d_grad,
        (hipStream_t)stream);
    if(e != hipSuccess) return fail(AGX_EDEVICE, std::string("eval launch: ") + hipGetErrorString(e));
    return AGX_OK;
}

int agx_eval_hess_many_device(const agx_potential* pot, long n, const double* d_pos, double t,
    double* d_phi, double* d_grad, double* d_hess, int flags, void* stream)
{
    (void)t;   // the Multipole potential is static
    if(!pot) return fail(AGX_EINVAL, "NULL potential");
    if(n < 0 || (n > 0 && !d_pos)) return fail(AGX_EINVAL, "invalid point array");
    if(flags == AGX_PRECISION_FP32) return fail(AGX_EUNSUPPORTED, "second derivatives are evaluated in FP64 only");
    if(flags != AGX_PRECISION_FP64) return fail(AGX_EINVAL, "invalid flags");
    hipError_t e = agxLaunchEvalHess(pot->ddev, n, d_pos, d_phi, d_grad, d_hess, (hipStream_t)stream);
    if(e != hipSuccess) return fail(AGX_EDEVICE, std::string("hessian launch: ") + hipGetErrorString(e));
    return AGX_OK;
}

}  // extern "C"

namespace {
// makes the potential's device current for the scope of a call and restores the caller's
struct DeviceScope {
    int prev = -1;
    explicit DeviceScope(int dev)
    {
        if(hipGetDevice(&prev) != hipSuccess) prev = -1;
        if(prev != dev) HIPCHECK(hipSetDevice(dev));
    }
    ~DeviceScope() { if(prev >= 0) (void)hipSetDevice(prev); }
};

// the handle's staging buffer with room for `doubles` values (pot->evalMutex held)
double* evalStaging(const agx_potential* pot, size_t doubles)
{
    if(!pot->evalStream) HIPCHECK(hipStreamCreateWithFlags(&pot->evalStream, hipStreamNonBlocking));
    if(pot->evalCap < doubles) {
        if(pot->evalBuf) HIPCHECK(hipFree(pot->evalBuf));
        pot->evalBuf = nullptr;
        pot->evalCap = 0;
        HIPCHECK(hipMalloc(&pot->evalBuf, doubles * sizeof(double)));
        pot->evalCap = doubles;
    }
    return pot->evalBuf;
}
}  // namespace

extern "C" {

int agx_eval_hess_many(const agx_potential* pot, long n, const double* pos, double t,
    double* phi, double* grad, double* hess, int flags)
{
    g_err.clear();
    if(!pot) return fail(AGX_EINVAL, "NULL potential");
    if(n == 0) return AGX_OK;
    if(n < 0 || !pos) return fail(AGX_EINVAL, "invalid point array");
    try {
        std::lock_guard<std::mutex> lock(pot->evalMutex);
        DeviceScope scope(pot->device);
        double* w = evalStaging(pot, (size_t)n * (3 + (phi ? 1 : 0) + (grad ? 3 : 0) + (hess ? 6 : 0)));
        size_t off = 0;
        auto take = [&](bool want, long count) { double* q = want ? w + off : nullptr; if(want) off += count; return q; };
        double* dp = take(true, n * 3);
        double* dphi = take(phi != nullptr, n);
        double* dg = take(grad != nullptr, n * 3);
        double* dh = take(hess != nullptr, n * 6);
        hipStream_t st = pot->evalStream;
        HIPCHECK(hipMemcpyAsync(dp, pos, n * 3 * 8, hipMemcpyHostToDevice, st));
        int rc = agx_eval_hess_many_device(pot, n, dp, t, dphi, dg, dh, flags, st);
        if(rc) return rc;
        if(phi) HIPCHECK(hipMemcpyAsync(phi, dphi, n * 8, hipMemcpyDeviceToHost, st));
        if(grad) HIPCHECK(hipMemcpyAsync(grad, dg, n * 3 * 8, hipMemcpyDeviceToHost, st));
        if(hess) HIPCHECK(hipMemcpyAsync(hess, dh, n * 6 * 8, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        return AGX_OK;
    } catch(std::exception& e) { return fail(AGX_EDEVICE, e.what()); }
}

int agx_eval_many(const agx_potential* pot, long n, const double* pos, double t,
    double* phi, double* grad, int flags)
{
    g_err.clear();
    if(!pot) return fail(AGX_EINVAL, "NULL potential");
    if(n == 0) return AGX_OK;
    if(n < 0 || !pos) return fail(AGX_EINVAL, "invalid point array");
    try {
        std::lock_guard<std::mutex> lock(pot->evalMutex);
        DeviceScope scope(pot->device);
        double* w = evalStaging(pot, (size_t)n * (3 + (phi ? 1 : 0) + (grad ? 3 : 0)));
        size_t off = 0;
        auto take = [&](bool want, long count) { double* q = want ? w + off : nullptr; if(want) off += count; return q; };
        double* dp = take(true, n * 3);
        double* dphi = take(phi != nullptr, n);
        double* dg = take(grad != nullptr, n * 3);
        hipStream_t st = pot->evalStream;
        HIPCHECK(hipMemcpyAsync(dp, pos, n * 3 * 8, hipMemcpyHostToDevice, st));
        int rc = agx_eval_many_device(pot, n, dp, t, dphi, dg, flags, st);
        if(rc) return rc;
        if(phi) HIPCHECK(hipMemcpyAsync(phi, dphi, n * 8, hipMemcpyDeviceToHost, st));
        if(grad) HIPCHECK(hipMemcpyAsync(grad, dg, n * 3 * 8, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        return AGX_OK;
    } catch(std::exception& e) { return fail(AGX_EDEVICE, e.what()); }
}

double agx_eval_potential(const agx_potential* pot, const double pos[3], double t,
    double deriv[3], double deriv2[6])
{
    double phi = NAN, g[3], h[6];
    if(deriv2) {   // {dx2, dy2, dz2, dxdy, dydz, dxdz} of Phi, as interface_c.cpp:81-88
        if(agx_eval_hess_many(pot, 1, pos, t, &phi, g, h, AGX_PRECISION_FP64) != AGX_OK) {
            for(int i = 0; i < 6; i++) deriv2[i] = NAN;
            return NAN;
        }
        for(int i = 0; i < 6; i++) deriv2[i] = h[i];
    } else if(agx_eval_many(pot, 1, pos, t, &phi, g, AGX_PRECISION_FP64) != AGX_OK)
        return NAN;
    if(deriv) for(int i = 0; i < 3; i++) deriv[i] = g[i];
    return phi;
}

int agx_set_orbit_launch(int num_blocks, int block_size)
{
    if(num_blocks < 0 || (block_size != 0 && block_size != agxOrbitKernelBlockSize()))
        return fail(AGX_EINVAL, "num_blocks must be >= 0 and block_size 0 or " +
            std::to_string(agxOrbitKernelBlockSize()));
    g_orbitBlocks = num_blocks;
    return AGX_OK;
}

int agx_set_fetch(int eval_mode, int orbit_mode)
{
    auto ok = [](int m) { return m == AGX_FETCH_SCALAR || m == AGX_FETCH_VEC2 || m == AGX_FETCH_COOP ||
                                 m == AGX_FETCH_CELL; };
    if(!ok(eval_mode) || !ok(orbit_mode)) return fail(AGX_EINVAL, "fetch mode must be 0 to 3");
    g_fetchEval = eval_mode;
    g_fetchOrbit = orbit_mode;
    return AGX_OK;
}

int agx_set_orbit_waves(int waves)
{
    if(waves != 1 && waves != 2) return fail(AGX_EINVAL, "waves must be 1 or 2");
    g_orbitWaves = waves;
    return AGX_OK;
}

int agx_set_orbit_groups(int stride, long orbits, long min_batch)
{
    if(stride < -1 || orbits < -1 || min_batch < 0) return fail(AGX_EINVAL, "invalid group-mode settings");
    g_groupStride = stride;
    g_groupOrbits = orbits;
    g_groupMinBatch = min_batch;
    return AGX_OK;
}

int agx_set_orbit_probe(int steps, long min_batch)
{
    if(steps < 0 || min_batch < 0) return fail(AGX_EINVAL, "invalid probe settings");
    g_probeSteps = steps;
    g_probeMinBatch = min_batch;
    return AGX_OK;
}

int agx_get_orbit_schedule(int* group_stride, long* group_orbits, long* group_min_batch, int* probe_steps,
    long* probe_min_batch)
{
    if(group_stride) *group_stride = g_groupStride;
    if(group_orbits) *group_orbits = g_groupOrbits;
    if(group_min_batch) *group_min_batch = g_groupMinBatch;
    if(probe_steps) *probe_steps = g_probeSteps;
    if(probe_min_batch) *probe_min_batch = g_probeMinBatch;
    return AGX_OK;
}

int agx_get_orbit_config(int* num_blocks, int* fetch_eval, int* fetch_orbit, int* waves)
{
    if(num_blocks) *num_blocks = g_orbitBlocks;
    if(fetch_eval) *fetch_eval = g_fetchEval;
    if(fetch_orbit) *fetch_orbit = g_fetchOrbit;
    if(waves) *waves = g_orbitWaves;
    return AGX_OK;
}

}  // extern "C"

// Lyapunov bookkeeping buffers of the variational kernel (device pointers; bins null: none)
struct LyapDev {
    const double* period = nullptr;
    const double* maxTime = nullptr;
    double* bins = nullptr;
    int* nbins = nullptr;
    double* lastTime = nullptr;
    int maxBins = 0;
};

// overflow pool of every-step recording (AgxOrbitArgs::poolCounter and the fields after it)
struct RecordPool {
    int* chunkDir = nullptr;
    unsigned int* counter = nullptr;
    long base = 0;
    int chunk = 0, maxChunks = 0, chunks = 0;
};

// bytes of the schedule's workspace for n orbits: agxLaunchSchedule's keys and sort storage,
// then the order itself (n ints)
static size_t scheduleWorkspaceBytes(long n)
{
    return agxScheduleWorkspace(n) + ((size_t)n * 4 + 255) / 256 * 256;
}

// group stride of large probed batches: the probe ranks the truly longest orbits first, and one
// wavefront in this many integrates them with a group of lanes per orbit
constexpr int PROBED_GROUP_STRIDE = 8;

// Re-rank the work order `ord` (agxLaunchSchedule's, inside `work`) from a probe: every orbit is
// integrated for at most g_probeSteps steps, without output but the time it reached, and the
// orbits are sorted by the steps they would take at that rate.  Returns the new order (inside
// `work`).
static int* probeOrder(const agx_potential* pot, long n, const double* d_ic, const double* d_tstart,
    const double* d_ttotal, double accuracy, double Omega, long maxNumSteps, int precision, int* ord, void* work,
    hipStream_t st, AgxOrbitArgs* dArgs)
{
    AgxOrbitArgs P{};
    P.n = n; P.ic = d_ic; P.tstart = d_tstart; P.ttotal = d_ttotal; P.trajsize = nullptr; P.trajOffset = nullptr;
    P.order = ord; P.accuracy = accuracy; P.Omega = Omega;
    P.maxNumSteps = std::min<long>(maxNumSteps, g_probeSteps);
    P.counter = pot->dcounter;
    P.outTend = agxScheduleProbeTimes(work, n);
    int blocks, bs;
    orbitGeometry(n, precision, pot->nfwMass != 0, blocks, bs);
    HIPCHECK(agxLaunchFillU32(pot->dcounter, 2, 0u, st));
    HIPCHECK(agxLaunchOrbit(pot->dev, pot->ddev, precision, g_fetchOrbit, g_orbitWaves, P, blocks, bs, st,
        dArgs ? dArgs : pot->dargs));
    int* fin = nullptr;
    HIPCHECK(agxLaunchScheduleFromProbe(n, d_tstart, d_ttotal, g_probeSteps, ord, work, &fin, st));
    return fin;
}

// Group mode of a DOP853 batch (agx_set_orbit_groups): sets A.groupOrbits, A.groupStride and
// A.groupCounter (the word after the batch's queue counter, cleared with it) and widens `blocks`
// so that the groups and the one-lane rest both fit.  probedOrder: the probe's work order (inside
// the batch's workspace, ours to permute) or null; after a probe the group wavefronts' orbits are
// dealt round-robin (stream st).  eligible: the batch's own conditions (no deviation vectors,
// targets or every-step recording).
static void setupGroups(const agx_potential* pot, long n, int precision, int flags, bool eligible,
    int* probedOrder, unsigned int* groupCounter, AgxOrbitArgs& A, int& blocks, int bs, hipStream_t st)
{
    const bool probed = probedOrder != nullptr;
    int groupStride = g_groupStride;
    if(groupStride < 0) {
        // automatic: the fewer orbits per lane, the more the longest orbits (rather than the
        // total work) set the batch's duration (measured on the uneven 1-300-period workload:
        // 131k orbits -11% at one wavefront in 4, 262k -4% at one in 8; 1M orbits after the
        // probe 7.5 -> 6.5 s at one in 8)
        const long lanes = (long)blocks * bs;
        groupStride = n * 8 <= lanes ? 1 : n <= 2 * lanes ? 4 : n <= 4 * lanes ? 8 : probed ? PROBED_GROUP_STRIDE : 0;
    }
    // DOP853: node records or per-cell tables at one wavefront per SIMD; DPRKN8: per-cell tables
    const bool rkn = (flags & AGX_METHOD_DPRKN8) != 0;
    const bool kernelOk = rkn ? g_fetchOrbit == AGX_FETCH_CELL && agxOrbitRKNCellsFit(pot->dev) :
        (g_fetchOrbit == AGX_FETCH_VEC2 || g_fetchOrbit == AGX_FETCH_CELL) && g_orbitWaves == 1;
    if(!(eligible && groupStride > 0 && n >= g_groupMinBatch && precision == AGX_PRECISION_FP64 && kernelOk &&
         pot->nfwMass == 0 && !pot->dev.interp1d))
        return;
    const long waves = (long)blocks * (bs / 64);
    const long stride = std::min<long>(groupStride, waves);   // wavefront 0 is always a group one
    long groupWaves = 0;
    for(int b = 0; b < blocks; b++)
        for(int w = 0; w < bs / 64; w++) groupWaves += agxIsGroupWave(b, w, bs / 64, (int)stride);
    A.groupOrbits = std::min<long>(n, g_groupOrbits >= 0 ? g_groupOrbits : groupWaves * (64 / 8));
    // with every wavefront in group mode no lane reads the single-lane queue: all the orbits then
    // go to the groups
    if(stride == 1 && A.groupOrbits > 0) A.groupOrbits = n;
    A.groupStride = A.groupOrbits > 0 ? (int)stride : 0;
    A.groupCounter = groupCounter;
    if(A.groupStride == 0) { A.groupOrbits = 0; return; }
    // one of the longest orbits per group wavefront (each takes eight at the start)
    if(probed && stride > 1 && A.groupOrbits == groupWaves * 8 && A.groupOrbits <= 8192)
        HIPCHECK(agxLaunchDealGroups(probedOrder, A.groupOrbits, st));
    // enough blocks for the group orbits (8 lanes each) and the rest (one lane each)
    int dev = 0, cus = 0;
    HIPCHECK(hipGetDevice(&dev));
    HIPCHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const long maxBlocks = g_orbitBlocks > 0 ? g_orbitBlocks : (long)cus * std::max(1, rkn ?
        agxOrbitKernelRKNBlocksPerCU(precision, false) :
        agxOrbitKernelBlocksPerCU(precision, g_fetchOrbit, g_orbitWaves, false, false));
    const long gw = A.groupStride * ((A.groupOrbits + 7) / 8);
    const long lw = A.groupStride > 1 ? ((n - A.groupOrbits + 63) / 64) * A.groupStride / (A.groupStride - 1) : 0;
    const long needWaves = std::max(gw, lw);
    blocks = (int)std::min(maxBlocks, std::max<long>(blocks, (needWaves + bs / 64 - 1) / (bs / 64)));
}

// the batch entry points' common body; d_dev non-null or ly.bins non-null: the variational
// kernel (der=True: six vectors; lyapunov only: one)
static int orbitBatchDevice(const agx_potential* pot, long n, const double* d_ic,
    const double* d_tstart, const double* d_ttotal, const int* d_trajsize,
    const long* d_traj_offset, const int* d_order, double accuracy, double Omega,
    long maxNumSteps, double* d_time, double* d_traj, double* d_final, double* d_dev, long* d_nsteps,
    long* d_nattempts, long* d_nevals, double* d_maxde, int* d_status, int flags, void* stream,
    const LyapDev& ly = LyapDev(), const RecordPool& pool = RecordPool(), const AgxTargetList* d_targets = nullptr,
    bool allowGroups = true, void* scheduleWork = nullptr)
{
    if(!pot) return fail(AGX_EINVAL, "NULL potential");
    if(n == 0) return AGX_OK;
    if(n < 0 || n > 0x7fffffffL || !d_ic || !d_ttotal || !d_trajsize)
        return fail(AGX_EINVAL, "invalid orbit arrays");
    if((d_time || d_traj) && !d_traj_offset) return fail(AGX_EINVAL, "trajectory output needs offsets");
    if(!(accuracy > 0)) return fail(AGX_EINVAL, "accuracy must be positive");
    if(!(Omega > -INFINITY && Omega < INFINITY)) return fail(AGX_EINVAL, "Omega must be finite");
    if(maxNumSteps <= 0) return fail(AGX_EINVAL, "maxNumSteps must be positive");
    const int precision = flags & AGX_PRECISION_MASK;
    if((precision != AGX_PRECISION_FP64 && precision != AGX_PRECISION_FP32) ||
        (flags & ~(AGX_PRECISION_MASK | AGX_ORDER_AS_GIVEN | AGX_METHOD_DPRKN8)))
        return fail(AGX_EINVAL, "invalid flags");
    const bool variational = d_dev || ly.bins;
    if(variational && (precision != AGX_PRECISION_FP64 || (flags & AGX_METHOD_DPRKN8)))
        return fail(AGX_EUNSUPPORTED, "deviation vectors are integrated with DOP853 in FP64 only");
    if(d_dev && !d_traj_offset) return fail(AGX_EINVAL, "deviation vectors need trajectory offsets");
    if(d_targets && (variational || precision != AGX_PRECISION_FP64))
        return fail(AGX_EUNSUPPORTED, "targets are collected in FP64, without deviation vectors");
    try {
        hipStream_t st = (hipStream_t)stream;
        void* work = nullptr;
        int* probedOrder = nullptr;   // the probe's order, inside the workspace (ours to permute)
        if(!d_order && !(flags & AGX_ORDER_AS_GIVEN)) {
            // longest expected integrations first (agx_schedule.hip); stream-ordered workspace,
            // or the caller's (scheduleWork: at least scheduleWorkspaceBytes(n) bytes)
            if(scheduleWork) work = scheduleWork;
            else HIPCHECK(hipMallocAsync(&work, scheduleWorkspaceBytes(n), st));
            int* ord = reinterpret_cast<int*>(static_cast<char*>(work) + agxScheduleWorkspace(n));
            hipError_t e = agxLaunchSchedule(pot->dev, pot->ddev, n, d_ic, d_ttotal, ord,
                work, st);
            if(e != hipSuccess) {
                if(!scheduleWork) (void)hipFreeAsync(work, st);
                return fail(AGX_EDEVICE, std::string("schedule: ") + hipGetErrorString(e));
            }
            d_order = ord;
            if(g_probeSteps > 0 && n >= g_probeMinBatch && !variational && !d_targets && !pool.counter)
                d_order = probedOrder = probeOrder(pot, n, d_ic, d_tstart, d_ttotal, accuracy, Omega, maxNumSteps,
                    precision, ord, work, st, nullptr);
        }
        AgxOrbitArgs A;
        A.n = n; A.ic = d_ic; A.tstart = d_tstart; A.ttotal = d_ttotal; A.trajsize = d_trajsize;
        A.trajOffset = d_traj_offset; A.order = d_order; A.accuracy = accuracy; A.Omega = Omega;
        A.maxNumSteps = maxNumSteps; A.outTime = d_time; A.outTraj = d_traj; A.outFinal = d_final;
        A.outNsteps = d_nsteps; A.outNattempts = d_nattempts; A.outNevals = d_nevals;
        A.outMaxde = d_maxde; A.outStatus = d_status; A.counter = pot->dcounter;
        if(pool.counter) {
            A.chunkDir = pool.chunkDir; A.poolCounter = pool.counter; A.poolBase = pool.base;
            A.poolChunk = pool.chunk; A.maxChunks = pool.maxChunks; A.poolChunks = pool.chunks;
        }
        A.targets = d_targets;
        int blocks, bs;
        orbitGeometry(n, precision, pot->nfwMass != 0, blocks, bs, flags & AGX_METHOD_DPRKN8, d_targets != nullptr);
        // group mode: the first orbits of the order (the longest expected) on groups of 8 lanes
        if(allowGroups)
            setupGroups(pot, n, precision, flags, !variational && !d_targets && !pool.counter, probedOrder,
                pot->dcounter + 1, A, blocks, bs, st);
        HIPCHECK(agxLaunchFillU32(pot->dcounter, 2, 0u, st));
        hipError_t e;
        if(variational) {
            // the kernel's LDS allows one block per CU; per-lane variational state in a
            // stream-ordered workspace
            int dev = 0, cus = 0;
            HIPCHECK(hipGetDevice(&dev));
            HIPCHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
            int vb = cus * std::max(1, agxOrbitVarKernelBlocksPerCU(pot->nfwMass != 0));
            const long need = (n + bs - 1) / bs;
            if(need < vb) vb = (int)std::max<long>(need, 1);
            void* vstate = nullptr;
            HIPCHECK(hipMallocAsync(&vstate, agxOrbitVarStateDoubles(vb) * sizeof(double), st));
            e = agxLaunchOrbitVar(pot->dev, pot->ddev, A, vb, static_cast<double*>(vstate), d_dev, ly.period,
                ly.maxTime, ly.bins, ly.nbins, ly.lastTime, ly.maxBins, st);
            HIPCHECK(hipFreeAsync(vstate, st));
        } else
            e = launchOrbitMethod(pot, precision, flags, A, blocks, bs, st);
        if(work && !scheduleWork) HIPCHECK(hipFreeAsync(work, st));
        if(e != hipSuccess) return fail(AGX_EDEVICE, std::string("orbit launch: ") + hipGetErrorString(e));
        return AGX_OK;
    } catch(std::exception& e) { return fail(AGX_EDEVICE, e.what()); }
}

extern "C" {

int agx_orbit_batch_device(const agx_potential* pot, long n, const double* d_ic,
    const double* d_tstart, const double* d_ttotal, const int* d_trajsize,
    const long* d_traj_offset, const int* d_order, double accuracy, double Omega,
    long maxNumSteps, double* d_time, double* d_traj, double* d_final, long* d_nsteps,
    long* d_nattempts, long* d_nevals, double* d_maxde, int* d_status, int flags, void* stream)
{
    return orbitBatchDevice(pot, n, d_ic, d_tstart, d_ttotal, d_trajsize, d_traj_offset, d_order, accuracy,
        Omega, maxNumSteps, d_time, d_traj, d_final, nullptr, d_nsteps, d_nattempts, d_nevals, d_maxde,
        d_status, flags, stream);
}

int agx_orbit_var_batch_device(const agx_potential* pot, long n, const double* d_ic,
    const double* d_tstart, const double* d_ttotal, const int* d_trajsize,
    const long* d_traj_offset, const int* d_order, double accuracy, double Omega,
    long maxNumSteps, double* d_time, double* d_traj, double* d_final, double* d_dev, long* d_nsteps,
    long* d_nattempts, long* d_nevals, double* d_maxde, int* d_status, int flags, void* stream)
{
    if(!d_dev) return fail(AGX_EINVAL, "NULL deviation-vector output");
    return orbitBatchDevice(pot, n, d_ic, d_tstart, d_ttotal, d_trajsize, d_traj_offset, d_order, accuracy,
        Omega, maxNumSteps, d_time, d_traj, d_final, d_dev, d_nsteps, d_nattempts, d_nevals, d_maxde,
        d_status, flags, stream);
}

}  // extern "C"

// ---- orbit batches without allocation (agx_orbit_batch_device_noalloc) ----
// The schedule's workspace (sort keys, the radix sort's storage and the work order) of at most
// NOALLOC_MAX_ORBITS orbits lives in this array: one copy per device, held until the process
// ends.  Host code never passes `scratch` itself to a kernel (on the host it names a stand-in,
// not device memory): it uses the device address taken once per device by scratchAddress().
// Nothing in it needs clearing between calls (the cost kernel writes every key and index it
// sorts).  Calls that overlap on one device race on it.
static constexpr long NOALLOC_MAX_ORBITS = 1L << 20;
static constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)48 << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratchMutex;
void* g_scratchPtr[MAX_DEVICES] = {};

// device address of `scratch` on the current device (hipGetSymbolAddress on the first call)
int scratchAddress(void** out)
{
    int dev = 0;
    if(hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES)
        return fail(AGX_EDEVICE, "no usable current HIP device");
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    if(!g_scratchPtr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if(e != hipSuccess || !p)
            return fail(AGX_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratchPtr[dev] = p;
    }
    *out = g_scratchPtr[dev];
    return AGX_OK;
}
}  // namespace

extern "C" {

int agx_orbit_batch_device_noalloc(const agx_potential* pot, long n, const double* d_ic,
    const double* d_tstart, const double* d_ttotal, const int* d_trajsize,
    const long* d_traj_offset, const int* d_order, double accuracy, double Omega,
    long maxNumSteps, double* d_time, double* d_traj, double* d_final, long* d_nsteps,
    long* d_nattempts, long* d_nevals, double* d_maxde, int* d_status, int flags, void* stream)
{
    g_err.clear();
    if(!pot) return fail(AGX_EINVAL, "NULL potential");
    if(n > NOALLOC_MAX_ORBITS)
        return fail(AGX_EINVAL, "agx_orbit_batch_device_noalloc: at most 1048576 orbits per call (got " +
            std::to_string(n) + ")");
    int dev = -1;
    if(hipGetDevice(&dev) != hipSuccess || dev != pot->device)
        return fail(AGX_EINVAL, "agx_orbit_batch_device_noalloc: the potential's device must be current");
    void* work = nullptr;
    if(n > 0 && !d_order && !(flags & AGX_ORDER_AS_GIVEN)) {
        if(scheduleWorkspaceBytes(n) > NOALLOC_SCRATCH_BYTES)
            return fail(AGX_EINVAL, "agx_orbit_batch_device_noalloc: the schedule of " + std::to_string(n) +
                " orbits needs more than the library's scratch memory");
        const int rc = scratchAddress(&work);
        if(rc) return rc;
    }
    return orbitBatchDevice(pot, n, d_ic, d_tstart, d_ttotal, d_trajsize, d_traj_offset, d_order, accuracy,
        Omega, maxNumSteps, d_time, d_traj, d_final, nullptr, d_nsteps, d_nattempts, d_nevals, d_maxde,
        d_status, flags, stream, LyapDev(), RecordPool(), nullptr, true, work);
}

}  // extern "C"

// one integration of a fixed batch captured as a HIP graph (counter fill + persistent kernel)
struct agx_orbit_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void* work = nullptr;          // work order (n ints) + private queue counter
    int device = 0;
    ~agx_orbit_graph()
    {
        if(exec) (void)hipGraphExecDestroy(exec);
        if(graph) (void)hipGraphDestroy(graph);
        if(work) (void)hipFree(work);
    }
};

extern "C" {

agx_orbit_graph* agx_orbit_graph_create(const agx_potential* pot, long n, const double* d_ic,
    const double* d_tstart, const double* d_ttotal, const int* d_trajsize, const long* d_traj_offset,
    double accuracy, double Omega, long maxNumSteps, double* d_time, double* d_traj, double* d_final,
    long* d_nsteps, long* d_nattempts, long* d_nevals, double* d_maxde, int* d_status, int flags)
{
    g_err.clear();
    if(!pot) { fail(AGX_EINVAL, "NULL potential"); return nullptr; }
    if(n <= 0 || n > 0x7fffffffL || !d_ic || !d_ttotal || !d_trajsize) { fail(AGX_EINVAL, "invalid orbit arrays"); return nullptr; }
    if((d_time || d_traj) && !d_traj_offset) { fail(AGX_EINVAL, "trajectory output needs offsets"); return nullptr; }
    if(!(accuracy > 0) || !(Omega > -INFINITY && Omega < INFINITY) || maxNumSteps <= 0) {
        fail(AGX_EINVAL, "invalid accuracy, Omega or maxNumSteps"); return nullptr;
    }
    const int precision = flags & AGX_PRECISION_MASK;
    if((precision != AGX_PRECISION_FP64 && precision != AGX_PRECISION_FP32) ||
        (flags & ~(AGX_PRECISION_MASK | AGX_ORDER_AS_GIVEN | AGX_METHOD_DPRKN8))) { fail(AGX_EINVAL, "invalid flags"); return nullptr; }
    try {
        std::unique_ptr<agx_orbit_graph> G(new agx_orbit_graph);
        HIPCHECK(hipSetDevice(pot->device));
        G->device = pot->device;
        // work order, queue counter, the kernel's argument block, the schedule's workspace
        const size_t ordBytes = ((size_t)n * 4 + 255) / 256 * 256, schedBytes = agxScheduleWorkspace(n),
            argBytes = align256(sizeof(AgxOrbitArgs));
        HIPCHECK(hipMalloc(&G->work, ordBytes + 256 + argBytes + schedBytes));
        char* w = static_cast<char*>(G->work);
        int* ord = reinterpret_cast<int*>(w);
        unsigned int* counter = reinterpret_cast<unsigned int*>(w + ordBytes);
        AgxOrbitArgs* dArgs = reinterpret_cast<AgxOrbitArgs*>(w + ordBytes + 256);
        const int* order = nullptr;
        int* probedOrder = nullptr;
        void* schedWork = w + ordBytes + 256 + argBytes;
        if(!(flags & AGX_ORDER_AS_GIVEN)) {   // the work order, computed once from the initial conditions
            HIPCHECK(agxLaunchSchedule(pot->dev, pot->ddev, n, d_ic, d_ttotal, ord, schedWork, nullptr));
            order = ord;
            // large batches: the order from the probe, as agx_orbit_batch_device takes it (the
            // probe's order stays in the graph's workspace)
            if(g_probeSteps > 0 && n >= g_probeMinBatch)
                order = probedOrder = probeOrder(pot, n, d_ic, d_tstart, d_ttotal, accuracy, Omega, maxNumSteps,
                    precision, ord, schedWork, nullptr, dArgs);
            HIPCHECK(hipDeviceSynchronize());
        }
        AgxOrbitArgs A;
        A.n = n; A.ic = d_ic; A.tstart = d_tstart; A.ttotal = d_ttotal; A.trajsize = d_trajsize;
        A.trajOffset = d_traj_offset; A.order = order; A.accuracy = accuracy; A.Omega = Omega;
        A.maxNumSteps = maxNumSteps; A.outTime = d_time; A.outTraj = d_traj; A.outFinal = d_final;
        A.outNsteps = d_nsteps; A.outNattempts = d_nattempts; A.outNevals = d_nevals;
        A.outMaxde = d_maxde; A.outStatus = d_status; A.counter = counter;
        int blocks, bs;
        orbitGeometry(n, precision, pot->nfwMass != 0, blocks, bs, flags & AGX_METHOD_DPRKN8);
        // group mode as in agx_orbit_batch_device (any dealing of the probe's order happens now)
        setupGroups(pot, n, precision, flags, true, probedOrder, counter + 1, A, blocks, bs, nullptr);
        HIPCHECK(hipDeviceSynchronize());
        hipStream_t st;
        HIPCHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        std::unique_ptr<void, void(*)(void*)> sg((void*)st, [](void* q) { (void)hipStreamDestroy((hipStream_t)q); });
        HIPCHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        hipError_t e = agxLaunchFillU32(counter, 2, 0u, st);   // queue counter and group counter
        if(e == hipSuccess) e = launchOrbitMethod(pot, precision, flags, A, blocks, bs, st, dArgs);
        HIPCHECK(hipStreamEndCapture(st, &G->graph));
        HIPCHECK(e);
        HIPCHECK(hipGraphInstantiate(&G->exec, G->graph, nullptr, nullptr, 0));
        return G.release();
    } catch(std::exception& e) { fail(AGX_EDEVICE, e.what()); return nullptr; }
}

int agx_orbit_graph_launch(agx_orbit_graph* g, void* stream)
{
    if(!g) return fail(AGX_EINVAL, "NULL graph");
    hipError_t e = hipGraphLaunch(g->exec, (hipStream_t)stream);
    return e == hipSuccess ? AGX_OK : fail(AGX_EDEVICE, hipGetErrorString(e));
}

void agx_orbit_graph_destroy(agx_orbit_graph* g) { delete g; }

}  // extern "C"

static int orbitBatchHost(const agx_potential* pot, long n, const double* ic, const double* tstart,
    const double* ttotal, const int* trajsize, double accuracy, double Omega, long maxNumSteps,
    double* out_time, double* out_traj, double* out_final, double* out_dev, long* out_nsteps,
    long* out_nattempts, long* out_nevals, double* out_maxde, int* out_status, int flags,
    double* out_lyap = nullptr, const double* maxTimeIn = nullptr)
{
    g_err.clear();
    if(!pot) return fail(AGX_EINVAL, "NULL potential");
    if(n == 0) return AGX_OK;
    if(n < 0 || !ic || !ttotal || !trajsize) return fail(AGX_EINVAL, "invalid orbit arrays");
    std::vector<long> offs(n);
    long rows = 0;
    for(long i = 0; i < n; i++) {
        if(trajsize[i] < 1) return fail(AGX_EUNSUPPORTED, "trajsize must be >= 1 (recording every "
            "timestep, trajsize=0, is not provided by the batch interface)");
        offs[i] = rows;
        rows += trajsize[i];
    }
    try {
        HIPCHECK(hipSetDevice(pot->device));
        DevBuf<double> dic(n * 6), dts(tstart ? n : 0), dtt(n), dtime(out_time ? rows : 0),
            dtraj(out_traj ? rows * 6 : 0), dfin(out_final ? n * 6 : 0), dmde(out_maxde ? n : 0),
            ddev(out_dev ? rows * 36 : 0);
        const bool lyap = out_lyap != nullptr;
        const int MAXB = 128;   // bins span at least 1% of the duration: at most 101 of them
        DevBuf<double> dper(lyap ? n : 0), dmaxt(lyap ? n : 0), dbins(lyap ? n * 2 * MAXB : 0),
            dlast(lyap ? n : 0);
        DevBuf<int> dnb(lyap ? n : 0);
        const bool needStatus = out_status || lyap;
        DevBuf<int> dsz(n), dst(needStatus ? n : 0);
        DevBuf<long> doff(n), dns(out_nsteps ? n : 0), dna(out_nattempts ? n : 0), dne(out_nevals ? n : 0);
        HIPCHECK(hipMemcpy(dic.p, ic, n * 48, hipMemcpyHostToDevice));
        if(tstart) HIPCHECK(hipMemcpy(dts.p, tstart, n * 8, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(dtt.p, ttotal, n * 8, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(dsz.p, trajsize, n * 4, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(doff.p, offs.data(), n * 8, hipMemcpyHostToDevice));
        LyapDev ly;
        if(lyap) {
            // bin width from the requested duration (or the caller's); orbit periods on the device
            std::vector<double> mt(n);
            for(long i = 0; i < n; i++) mt[i] = maxTimeIn ? maxTimeIn[i] : fabs(ttotal[i]);
            HIPCHECK(hipMemcpy(dmaxt.p, mt.data(), n * 8, hipMemcpyHostToDevice));
            HIPCHECK(agxLaunchTcirc(pot->ddev, n, dic.p, dts.p, Omega, dper.p, nullptr));
            ly.period = dper.p; ly.maxTime = dmaxt.p; ly.bins = dbins.p; ly.nbins = dnb.p;
            ly.lastTime = dlast.p; ly.maxBins = MAXB;
        }
        int rc = orbitBatchDevice(pot, n, dic.p, dts.p, dtt.p, dsz.p, doff.p, nullptr,
            accuracy, Omega, maxNumSteps, dtime.p, dtraj.p, dfin.p, ddev.p, dns.p, dna.p, dne.p, dmde.p,
            dst.p, flags, nullptr, ly);
        if(rc) return rc;
        HIPCHECK(hipDeviceSynchronize());
        if(lyap) {
            std::vector<double> per(n), bins((size_t)n * 2 * MAXB), last(n);
            std::vector<int> nb(n), stat(n);
            HIPCHECK(hipMemcpy(per.data(), dper.p, n * 8, hipMemcpyDeviceToHost));
            HIPCHECK(hipMemcpy(bins.data(), dbins.p, (size_t)n * 2 * MAXB * 8, hipMemcpyDeviceToHost));
            HIPCHECK(hipMemcpy(last.data(), dlast.p, n * 8, hipMemcpyDeviceToHost));
            HIPCHECK(hipMemcpy(nb.data(), dnb.p, n * 4, hipMemcpyDeviceToHost));
            HIPCHECK(hipMemcpy(stat.data(), dst.p, n * 4, hipMemcpyDeviceToHost));
            // orbits that stopped early were binned with the requested duration: integrate them
            // again with the duration they reached (the bin width is 1% of it)
            std::vector<long> redo;
            if(!maxTimeIn)
                for(long i = 0; i < n; i++)
                    if(stat[i] != 0 && last[i] == last[i] && last[i] != fabs(ttotal[i])) redo.push_back(i);
            std::vector<double> redoLyap(redo.size() * 2);
            if(!redo.empty()) {
                const long m = (long)redo.size();
                std::vector<double> ic2(m * 6), ts2(tstart ? m : 0), tt2(m), mt2(m);
                std::vector<int> sz2(m, 1);
                for(long j = 0; j < m; j++) {
                    const long i = redo[j];
                    for(int c = 0; c < 6; c++) ic2[j * 6 + c] = ic[i * 6 + c];
                    if(tstart) ts2[j] = tstart[i];
                    tt2[j] = ttotal[i];
                    mt2[j] = last[i];
                }
                // the same form of the variational equation (six vectors when der=True)
                std::vector<double> dev2(out_dev ? m * 36 : 0);
                int rc2 = orbitBatchHost(pot, m, ic2.data(), tstart ? ts2.data() : nullptr, tt2.data(), sz2.data(),
                    accuracy, Omega, maxNumSteps, nullptr, nullptr, nullptr, out_dev ? dev2.data() : nullptr,
                    nullptr, nullptr, nullptr, nullptr, nullptr, flags, redoLyap.data(), mt2.data());
                if(rc2) return rc2;
            }
            for(long i = 0; i < n; i++) {
                const double* b = bins.data() + (size_t)i * 2 * MAXB;
                agx::lyapunovFromBins(b, b + MAXB, std::min(nb[i], MAXB), last[i], per[i], out_lyap + i * 2);
            }
            for(size_t j = 0; j < redo.size(); j++) {
                out_lyap[redo[j] * 2] = redoLyap[j * 2];
                out_lyap[redo[j] * 2 + 1] = redoLyap[j * 2 + 1];
            }
        }
        if(out_time) HIPCHECK(hipMemcpy(out_time, dtime.p, rows * 8, hipMemcpyDeviceToHost));
        if(out_traj) HIPCHECK(hipMemcpy(out_traj, dtraj.p, rows * 48, hipMemcpyDeviceToHost));
        if(out_final) HIPCHECK(hipMemcpy(out_final, dfin.p, n * 48, hipMemcpyDeviceToHost));
        if(out_nsteps) HIPCHECK(hipMemcpy(out_nsteps, dns.p, n * 8, hipMemcpyDeviceToHost));
        if(out_nattempts) HIPCHECK(hipMemcpy(out_nattempts, dna.p, n * 8, hipMemcpyDeviceToHost));
        if(out_nevals) HIPCHECK(hipMemcpy(out_nevals, dne.p, n * 8, hipMemcpyDeviceToHost));
        if(out_maxde) HIPCHECK(hipMemcpy(out_maxde, dmde.p, n * 8, hipMemcpyDeviceToHost));
        if(out_status) HIPCHECK(hipMemcpy(out_status, dst.p, n * 4, hipMemcpyDeviceToHost));
        if(out_dev) HIPCHECK(hipMemcpy(out_dev, ddev.p, rows * 36 * 8, hipMemcpyDeviceToHost));
        return AGX_OK;
    } catch(std::exception& e) { return fail(AGX_EDEVICE, e.what()); }
}

extern "C" {

int agx_orbit_batch(const agx_potential* pot, long n, const double* ic, const double* tstart,
    const double* ttotal, const int* trajsize, double accuracy, double Omega, long maxNumSteps,
    double* out_time, double* out_traj, double* out_final, long* out_nsteps,
    long* out_nattempts, long* out_nevals, double* out_maxde, int* out_status, int flags)
{
    return orbitBatchHost(pot, n, ic, tstart, ttotal, trajsize, accuracy, Omega, maxNumSteps, out_time,
        out_traj, out_final, nullptr, out_nsteps, out_nattempts, out_nevals, out_maxde, out_status, flags);
}

int agx_lyapunov_fit(const double* t, const double* lnw, int nbins, double max_time, double period,
    double* out)
{
    g_err.clear();
    if(!out || (nbins > 0 && (!t || !lnw))) return fail(AGX_EINVAL, "NULL argument");
    agx::lyapunovFromBins(t, lnw, nbins, max_time, period, out);
    return AGX_OK;
}

int agx_orbit_var_batch(const agx_potential* pot, long n, const double* ic, const double* tstart,
    const double* ttotal, const int* trajsize, double accuracy, double Omega, long maxNumSteps,
    double* out_time, double* out_traj, double* out_final, double* out_dev, double* out_lyap,
    long* out_nsteps, long* out_nattempts, long* out_nevals, double* out_maxde, int* out_status, int flags)
{
    if(!out_dev && !out_lyap) return fail(AGX_EINVAL, "neither deviation vectors nor Lyapunov exponents requested");
    return orbitBatchHost(pot, n, ic, tstart, ttotal, trajsize, accuracy, Omega, maxNumSteps, out_time,
        out_traj, out_final, out_dev, out_nsteps, out_nattempts, out_nevals, out_maxde, out_status, flags,
        out_lyap);
}

}  // extern "C"

// ---------------- every-step recording and interpolated orbits ----------------
// The trajectory (and deviation vectors) of a batch kept in device memory: orbits with
// trajsize 0 record the start and the end of every step (orbit.cpp:38-45), the others trajsize
// rows at regular intervals.  The number of steps is known only at the end, so the first
// integration reserves a fixed block per every-step orbit, and an orbit that took more steps
// is integrated again with a block of its exact size (the integration does not depend on what
// is recorded, so the second one repeats the first step for step).  The quintic splines of
// agama.Orbit objects are built from these rows on first use (agx_orbit_interp.hip).
struct agx_orbit_record {
    int device = 0;
    long n = 0, deviceRows = 0;
    double Omega = 0;
    bool der = false;
    std::vector<long> off, cap, rows;   // per orbit: first device row, rows reserved, rows returned
    std::unique_ptr<DevBuf<double>> time, traj, dev;
    // interpolators: series 0..n-1 are the trajectories, n + 6*i + k the k-th deviation vector
    // of orbit i
    std::unique_ptr<DevBuf<AgxSeries>> series;
    std::unique_ptr<DevBuf<double>> nodes;
    std::unique_ptr<DevBuf<int>> count;
    std::vector<long> nodeOff;
    std::vector<int> hcount, hrev;
    long numSeries() const { return n * (der ? 7 : 1); }
};

namespace {

long g_recordRows = 0;   // first-pass row budget of agx_orbit_record_batch (0: default)

// rows an every-step orbit recorded: the start and each step's end; a zero-length (or
// non-finite) request records the start twice; a failure before the first step records nothing
long everyStepRows(long nsteps, double tstart, double ttotal)
{
    if(nsteps > 0) return nsteps + 1;
    const double te = tstart + ttotal;
    return ttotal == 0 || !(te > -INFINITY && te < INFINITY) ? 2 : 0;
}

void buildSplines(agx_orbit_record& R)
{
    if(R.series) return;
    const long ns = R.numSeries();
    std::vector<AgxSeries> S(ns);
    R.nodeOff.assign(ns, 0);
    long nodes = 0;
    for(long s = 0; s < ns; s++) {
        const long i = s < R.n ? s : (s - R.n) / 6, k = s < R.n ? -1 : (s - R.n) % 6;
        S[s].t = R.time->p + R.off[i];
        S[s].v = k < 0 ? R.traj->p + R.off[i] * 6 : R.dev->p + R.off[i] * 36 + k * R.cap[i] * 6;
        S[s].rows = R.rows[i];
        S[s].node = nodes;
        R.nodeOff[s] = nodes;
        nodes += R.rows[i];
    }
    std::unique_ptr<DevBuf<AgxSeries>> dS(new DevBuf<AgxSeries>(std::max<long>(ns, 1)));
    std::unique_ptr<DevBuf<double>> dN(new DevBuf<double>(std::max<long>(nodes, 1) * AGX_NODE_DOUBLES));
    std::unique_ptr<DevBuf<int>> dC(new DevBuf<int>(std::max<long>(ns, 1)));
    DevBuf<double> piv(std::max<long>(nodes, 1));
    DevBuf<int> rev(std::max<long>(ns, 1));
    if(ns) HIPCHECK(hipMemcpy(dS->p, S.data(), ns * sizeof(AgxSeries), hipMemcpyHostToDevice));
    HIPCHECK(agxLaunchBuildSplines(dS->p, ns, R.Omega, dN->p, piv.p, dC->p, rev.p, nullptr));
    R.hcount.assign(ns, 0);
    R.hrev.assign(ns, 0);
    if(ns) {
        HIPCHECK(hipMemcpy(R.hcount.data(), dC->p, ns * 4, hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(R.hrev.data(), rev.p, ns * 4, hipMemcpyDeviceToHost));
    }
    R.series = std::move(dS);
    R.nodes = std::move(dN);
    R.count = std::move(dC);
}

}  // namespace

extern "C" {

agx_orbit_record* agx_orbit_record_batch(const agx_potential* pot, long n, const double* ic,
    const double* tstart, const double* ttotal, const int* trajsize, double accuracy, double Omega,
    long maxNumSteps, int der, long* out_rows, double* out_final, long* out_nsteps, long* out_nattempts,
    long* out_nevals, double* out_maxde, int* out_status, int flags)
{
    g_err.clear();
    if(!pot) { fail(AGX_EINVAL, "NULL potential"); return nullptr; }
    if(n < 0 || n > 0x7fffffffL || (n > 0 && (!ic || !ttotal || !trajsize))) {
        fail(AGX_EINVAL, "invalid orbit arrays");
        return nullptr;
    }
    for(long i = 0; i < n; i++)
        if(trajsize[i] < 0) { fail(AGX_EINVAL, "trajsize must be >= 0"); return nullptr; }
    if(maxNumSteps <= 0) { fail(AGX_EINVAL, "maxNumSteps must be positive"); return nullptr; }
    try {
        HIPCHECK(hipSetDevice(pot->device));
        std::unique_ptr<agx_orbit_record> R(new agx_orbit_record);
        R->device = pot->device;
        R->n = n;
        R->Omega = Omega;
        R->der = der != 0;
        R->off.assign(n, 0); R->cap.assign(n, 0); R->rows.assign(n, 0);
        if(n == 0) {
            R->time.reset(new DevBuf<double>(1));
            R->traj.reset(new DevBuf<double>(6));
            R->dev.reset(new DevBuf<double>(R->der ? 36 : 0));
            return R.release();
        }
        long nEvery = 0;
        for(long i = 0; i < n; i++) nEvery += trajsize[i] == 0;
        // Rows for every-step orbits: a budget of 2^28 rows (15 GB; 2^25 with deviation vectors,
        // 11.5 GB; or the budget set by agx_set_record_rows), half of it as a block per orbit
        // (64 to 32768 rows, at most the step limit + 1) and as much again (at most the rest of
        // the budget) as an overflow pool of chunks of the block's size (at most 1024 rows) that
        // lanes take as their orbits outgrow the block, up to 64 chunks per orbit.
        const long budget = g_recordRows > 0 ? g_recordRows : R->der ? (1L << 25) : (1L << 28);
        long cap0 = nEvery ? std::max(g_recordRows > 0 ? 2L : 64L, std::min(budget / (2 * nEvery), 1L << 15)) : 0;
        cap0 = std::min(cap0, std::max(2L, maxNumSteps + 1));
        const long CH = std::max(4L, std::min(cap0, 1024L));
        const int MAXCH = 64;
        std::vector<int> ksz(n);
        long blocks = 0;
        for(long i = 0; i < n; i++) {
            const long c = trajsize[i] > 0 ? trajsize[i] : cap0;
            R->cap[i] = c;
            R->off[i] = blocks;
            blocks += c;
            ksz[i] = trajsize[i] > 0 ? trajsize[i] : -(int)c;
        }
        const long poolRows = nEvery == 0 ? 0 : g_recordRows > 0 ? budget - nEvery * cap0 :
            std::min(budget - nEvery * cap0, nEvery * cap0);
        const long poolChunks = std::max(0L, poolRows / CH);
        const long rows1 = blocks + poolChunks * CH;
        std::vector<double> fin(n * 6), mde(n);
        std::vector<long> ns(n), na(n), ne(n);
        std::vector<int> st(n);
        // one integration of the orbits idx (all when empty) into the blocks at R->off with
        // capacities ksz, writing into time/traj/dev
        auto integrate = [&](const std::vector<long>& idx, double* dtime, double* dtraj, double* ddev,
                             const RecordPool& pool) -> int {
            const long m = idx.empty() ? n : (long)idx.size();
            auto at = [&](long j) { return idx.empty() ? j : idx[j]; };
            std::vector<double> hic(m * 6), hts(tstart ? m : 0), htt(m);
            std::vector<int> hsz(m);
            std::vector<long> hoff(m);
            for(long j = 0; j < m; j++) {
                const long i = at(j);
                for(int c = 0; c < 6; c++) hic[j * 6 + c] = ic[i * 6 + c];
                if(tstart) hts[j] = tstart[i];
                htt[j] = ttotal[i];
                hsz[j] = ksz[i];
                hoff[j] = R->off[i];
            }
            DevBuf<double> dic(m * 6), dts(tstart ? m : 0), dtt(m), dfin(m * 6), dmde(m);
            DevBuf<int> dsz(m), dst(m);
            DevBuf<long> doff(m), dns(m), dna(m), dne(m);
            HIPCHECK(hipMemcpy(dic.p, hic.data(), m * 48, hipMemcpyHostToDevice));
            if(tstart) HIPCHECK(hipMemcpy(dts.p, hts.data(), m * 8, hipMemcpyHostToDevice));
            HIPCHECK(hipMemcpy(dtt.p, htt.data(), m * 8, hipMemcpyHostToDevice));
            HIPCHECK(hipMemcpy(dsz.p, hsz.data(), m * 4, hipMemcpyHostToDevice));
            HIPCHECK(hipMemcpy(doff.p, hoff.data(), m * 8, hipMemcpyHostToDevice));
            int rc = orbitBatchDevice(pot, m, dic.p, dts.p, dtt.p, dsz.p, doff.p, nullptr, accuracy, Omega,
                maxNumSteps, dtime, dtraj, dfin.p, ddev, dns.p, dna.p, dne.p, dmde.p, dst.p, flags, nullptr,
                LyapDev(), pool, nullptr, false);   // one-lane form in both passes: the same steps
            if(rc) return rc;
            HIPCHECK(hipDeviceSynchronize());
            std::vector<double> f2(m * 6), d2(m);
            std::vector<long> s2(m), a2(m), e2(m);
            std::vector<int> t2(m);
            HIPCHECK(hipMemcpy(f2.data(), dfin.p, m * 48, hipMemcpyDeviceToHost));
            HIPCHECK(hipMemcpy(d2.data(), dmde.p, m * 8, hipMemcpyDeviceToHost));
            HIPCHECK(hipMemcpy(s2.data(), dns.p, m * 8, hipMemcpyDeviceToHost));
            HIPCHECK(hipMemcpy(a2.data(), dna.p, m * 8, hipMemcpyDeviceToHost));
            HIPCHECK(hipMemcpy(e2.data(), dne.p, m * 8, hipMemcpyDeviceToHost));
            HIPCHECK(hipMemcpy(t2.data(), dst.p, m * 4, hipMemcpyDeviceToHost));
            for(long j = 0; j < m; j++) {
                const long i = at(j);
                for(int c = 0; c < 6; c++) fin[i * 6 + c] = f2[j * 6 + c];
                mde[i] = d2[j]; ns[i] = s2[j]; na[i] = a2[j]; ne[i] = e2[j]; st[i] = t2[j];
            }
            return AGX_OK;
        };
        // first pass: blocks + pool
        DevBuf<double> t1(rows1), x1(rows1 * 6), v1(R->der ? rows1 * 36 : 0);
        DevBuf<int> dir(nEvery ? n * MAXCH : 0);
        DevBuf<unsigned int> counter(1);
        RecordPool pool;
        if(nEvery) {
            HIPCHECK(agxLaunchFillU32(reinterpret_cast<unsigned int*>(dir.p), n * MAXCH, 0xFFFFFFFFu, nullptr));
            HIPCHECK(agxLaunchFillU32(counter.p, 1, 0u, nullptr));
            pool.chunkDir = dir.p; pool.counter = counter.p; pool.base = blocks; pool.chunk = (int)CH;
            pool.maxChunks = MAXCH; pool.chunks = (int)std::min<long>(poolChunks, 0x7fffffffL);
        }
        if(integrate({}, t1.p, x1.p, R->der ? v1.p : nullptr, pool)) return nullptr;
        std::vector<int> hdir;
        std::vector<long> need(n, 0), redo;
        for(long i = 0; i < n; i++)
            need[i] = trajsize[i] > 0 ? trajsize[i] : everyStepRows(ns[i], tstart ? tstart[i] : 0, ttotal[i]);
        for(long i = 0; i < n; i++) {
            if(trajsize[i] > 0 || need[i] <= R->cap[i]) continue;
            if(hdir.empty()) {
                hdir.resize(n * MAXCH);
                HIPCHECK(hipMemcpy(hdir.data(), dir.p, n * MAXCH * 4, hipMemcpyDeviceToHost));
            }
            const long chunks = (need[i] - R->cap[i] + CH - 1) / CH;
            bool ok = chunks <= MAXCH;
            for(long k = 0; ok && k < chunks; k++) ok = hdir[i * MAXCH + k] >= 0;
            if(!ok) redo.push_back(i);
        }
        // orbits the pool could not hold: integrated again with blocks of their size
        std::vector<long> off2(n, -1);
        long rows2 = 0;
        for(long i : redo) { off2[i] = rows2; rows2 += need[i] + 16; }
        DevBuf<double> t2(rows2), x2(rows2 * 6), v2(R->der ? rows2 * 36 : 0);
        if(!redo.empty()) {
            std::vector<long> off1 = R->off;
            for(long i : redo) { R->off[i] = off2[i]; ksz[i] = -(int)(need[i] + 16); }
            if(integrate(redo, t2.p, x2.p, R->der ? v2.p : nullptr, RecordPool())) return nullptr;
            for(long i : redo) {
                if(everyStepRows(ns[i], tstart ? tstart[i] : 0, ttotal[i]) != need[i]) {
                    fail(AGX_EDEVICE, "an orbit took a different number of steps when integrated again");
                    return nullptr;
                }
            }
            R->off = off1;
        }
        // gather every orbit's rows into one compact buffer: segments from the first pass
        // (block, then its pool chunks in order) or from the second
        std::vector<long> dstOff(n);
        long total = 0;
        for(long i = 0; i < n; i++) { dstOff[i] = total; total += need[i]; }
        struct Seg { std::vector<long> src, dst, cnt; };
        Seg a1, a6, ad1, a2s, a26, ad2;   // time/traj rows and dev rows, first and second pass
        for(long i = 0; i < n; i++) {
            const long r = need[i];
            if(off2[i] >= 0) {
                a2s.src.push_back(off2[i]); a2s.dst.push_back(dstOff[i]); a2s.cnt.push_back(r);
                for(int k = 0; R->der && k < 6; k++) {
                    ad2.src.push_back(off2[i] * 6 + k * (need[i] + 16));
                    ad2.dst.push_back(dstOff[i] * 6 + k * r); ad2.cnt.push_back(r);
                }
                continue;
            }
            const long inBlock = std::min(r, R->cap[i]);
            a1.src.push_back(R->off[i]); a1.dst.push_back(dstOff[i]); a1.cnt.push_back(inBlock);
            for(int k = 0; R->der && k < 6; k++) {
                ad1.src.push_back(R->off[i] * 6 + k * R->cap[i]);
                ad1.dst.push_back(dstOff[i] * 6 + k * r); ad1.cnt.push_back(inBlock);
            }
            for(long c = 0, left = r - inBlock; left > 0; c++, left -= CH) {
                const long chunk = hdir[i * MAXCH + c], len = std::min<long>(left, CH);
                const long at = inBlock + c * CH;
                a1.src.push_back(blocks + chunk * CH); a1.dst.push_back(dstOff[i] + at); a1.cnt.push_back(len);
                for(int k = 0; R->der && k < 6; k++) {
                    ad1.src.push_back((blocks + chunk * CH) * 6 + k * CH);
                    ad1.dst.push_back(dstOff[i] * 6 + k * r + at); ad1.cnt.push_back(len);
                }
            }
        }
        std::unique_ptr<DevBuf<double>> ct(new DevBuf<double>(std::max<long>(total, 1))),
            cx(new DevBuf<double>(std::max<long>(total, 1) * 6)),
            cv(new DevBuf<double>(R->der ? std::max<long>(total, 1) * 36 : 0));
        auto gather = [&](const Seg& g, int width, const double* src, double* dst) {
            const long k = (long)g.src.size();
            if(!k) return;
            DevBuf<long> ds(k), dd(k), dc(k);
            HIPCHECK(hipMemcpy(ds.p, g.src.data(), k * 8, hipMemcpyHostToDevice));
            HIPCHECK(hipMemcpy(dd.p, g.dst.data(), k * 8, hipMemcpyHostToDevice));
            HIPCHECK(hipMemcpy(dc.p, g.cnt.data(), k * 8, hipMemcpyHostToDevice));
            HIPCHECK(agxLaunchGatherRows(k, ds.p, dd.p, dc.p, width, src, dst, nullptr));
            HIPCHECK(hipDeviceSynchronize());
        };
        gather(a1, 1, t1.p, ct->p); gather(a1, 6, x1.p, cx->p);
        gather(a2s, 1, t2.p, ct->p); gather(a2s, 6, x2.p, cx->p);
        if(R->der) { gather(ad1, 6, v1.p, cv->p); gather(ad2, 6, v2.p, cv->p); }
        R->time = std::move(ct); R->traj = std::move(cx); R->dev = std::move(cv);
        R->deviceRows = total;
        for(long i = 0; i < n; i++) { R->off[i] = dstOff[i]; R->cap[i] = need[i]; R->rows[i] = need[i]; }
        for(long i = 0; i < n; i++) {
            if(out_rows) out_rows[i] = R->rows[i];
            if(out_final) for(int c = 0; c < 6; c++) out_final[i * 6 + c] = fin[i * 6 + c];
            if(out_nsteps) out_nsteps[i] = ns[i];
            if(out_nattempts) out_nattempts[i] = na[i];
            if(out_nevals) out_nevals[i] = ne[i];
            if(out_maxde) out_maxde[i] = mde[i];
            if(out_status) out_status[i] = st[i];
        }
        return R.release();
    } catch(std::exception& e) { fail(AGX_EDEVICE, e.what()); return nullptr; }
}

int agx_orbit_record_fetch(const agx_orbit_record* R, double* out_time, double* out_traj, double* out_dev)
{
    g_err.clear();
    if(!R) return fail(AGX_EINVAL, "NULL record");
    if(out_dev && !R->der) return fail(AGX_EINVAL, "the record holds no deviation vectors");
    try {
        HIPCHECK(hipSetDevice(R->device));
        // the blocks hold unused rows (reserved, or left by a first pass): gather the recorded
        // rows into one contiguous buffer on the device, then copy that
        const long n = R->n;
        long total = 0;
        for(long i = 0; i < n; i++) total += R->rows[i];
        if(total == 0) return AGX_OK;
        const long nseg = out_dev ? 7 * n : n;
        std::vector<long> src(nseg), dst(nseg), cnt(nseg);
        long o = 0;
        for(long i = 0; i < n; i++) {
            src[i] = R->off[i]; dst[i] = o; cnt[i] = R->rows[i];
            if(out_dev)   // six vectors of rows[i] rows each; the device block has cap[i] per vector
                for(int k = 0; k < 6; k++) {
                    src[n + i * 6 + k] = R->off[i] * 6 + k * R->cap[i];
                    dst[n + i * 6 + k] = o * 6 + k * R->rows[i];
                    cnt[n + i * 6 + k] = R->rows[i];
                }
            o += R->rows[i];
        }
        DevBuf<long> dsrc(nseg), ddst(nseg), dcnt(nseg);
        HIPCHECK(hipMemcpy(dsrc.p, src.data(), nseg * 8, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(ddst.p, dst.data(), nseg * 8, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(dcnt.p, cnt.data(), nseg * 8, hipMemcpyHostToDevice));
        DevBuf<double> buf(out_dev ? total * 36 : total * 6);
        if(out_time) {
            HIPCHECK(agxLaunchGatherRows(n, dsrc.p, ddst.p, dcnt.p, 1, R->time->p, buf.p, nullptr));
            HIPCHECK(hipMemcpy(out_time, buf.p, total * 8, hipMemcpyDeviceToHost));
        }
        if(out_traj) {
            HIPCHECK(agxLaunchGatherRows(n, dsrc.p, ddst.p, dcnt.p, 6, R->traj->p, buf.p, nullptr));
            HIPCHECK(hipMemcpy(out_traj, buf.p, total * 48, hipMemcpyDeviceToHost));
        }
        if(out_dev) {
            HIPCHECK(agxLaunchGatherRows(6 * n, dsrc.p + n, ddst.p + n, dcnt.p + n, 6, R->dev->p, buf.p, nullptr));
            HIPCHECK(hipMemcpy(out_dev, buf.p, total * 288, hipMemcpyDeviceToHost));
        }
        return AGX_OK;
    } catch(std::exception& e) { return fail(AGX_EDEVICE, e.what()); }
}

long agx_orbit_record_num_series(const agx_orbit_record* R) { return R ? R->numSeries() : -1; }

int agx_set_record_rows(long rows)
{
    if(rows < 0) return fail(AGX_EINVAL, "rows must be >= 0");
    g_recordRows = rows;
    return AGX_OK;
}

int agx_orbit_record_series_info(agx_orbit_record* R, long* out_nodes, int* out_reversed)
{
    g_err.clear();
    if(!R) return fail(AGX_EINVAL, "NULL record");
    try {
        HIPCHECK(hipSetDevice(R->device));
        buildSplines(*R);
        for(long s = 0; s < R->numSeries(); s++) {
            if(out_nodes) out_nodes[s] = R->hcount[s];
            if(out_reversed) out_reversed[s] = R->hrev[s];
        }
        return AGX_OK;
    } catch(std::exception& e) { return fail(AGX_EDEVICE, e.what()); }
}

int agx_orbit_record_node_times(agx_orbit_record* R, long series, double* out_t)
{
    g_err.clear();
    if(!R || !out_t) return fail(AGX_EINVAL, "NULL argument");
    if(series < 0 || series >= R->numSeries()) return fail(AGX_EINVAL, "series index out of range");
    try {
        HIPCHECK(hipSetDevice(R->device));
        buildSplines(*R);
        const long k = R->hcount[series];
        if(k) HIPCHECK(hipMemcpy2D(out_t, 8, R->nodes->p + R->nodeOff[series] * AGX_NODE_DOUBLES,
            AGX_NODE_DOUBLES * 8, 8, k, hipMemcpyDeviceToHost));
        return AGX_OK;
    } catch(std::exception& e) { return fail(AGX_EDEVICE, e.what()); }
}

int agx_orbit_record_nodes(agx_orbit_record* R, long series, double* out)
{
    g_err.clear();
    if(!R || !out) return fail(AGX_EINVAL, "NULL argument");
    if(series < 0 || series >= R->numSeries()) return fail(AGX_EINVAL, "series index out of range");
    try {
        HIPCHECK(hipSetDevice(R->device));
        buildSplines(*R);
        const long k = R->hcount[series];
        if(k) HIPCHECK(hipMemcpy(out, R->nodes->p + R->nodeOff[series] * AGX_NODE_DOUBLES,
            (size_t)k * AGX_NODE_DOUBLES * 8, hipMemcpyDeviceToHost));
        return AGX_OK;
    } catch(std::exception& e) { return fail(AGX_EDEVICE, e.what()); }
}

int agx_orbit_record_interp(agx_orbit_record* R, long m, const long* series, const double* t, int mode,
    double* out)
{
    g_err.clear();
    if(!R) return fail(AGX_EINVAL, "NULL record");
    if(m < 0 || (m > 0 && (!series || !t || !out))) return fail(AGX_EINVAL, "invalid query arrays");
    if(mode != 0 && mode != 1) return fail(AGX_EINVAL, "mode must be 0 or 1");
    const long ns = R->numSeries();
    for(long q = 0; q < m; q++)
        if(series[q] < 0 || series[q] >= ns) return fail(AGX_EINVAL, "series index out of range");
    if(m == 0) return AGX_OK;
    try {
        HIPCHECK(hipSetDevice(R->device));
        buildSplines(*R);
        const int width = mode == 0 ? 6 : 9;
        DevBuf<long> dq(m);
        DevBuf<double> dt(m), dout(m * width);
        HIPCHECK(hipMemcpy(dq.p, series, m * 8, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(dt.p, t, m * 8, hipMemcpyHostToDevice));
        HIPCHECK(agxLaunchEvalSplines(R->series->p, R->count->p, R->nodes->p, R->Omega, m, dq.p, dt.p, mode,
            dout.p, nullptr));
        HIPCHECK(hipMemcpy(out, dout.p, m * width * 8, hipMemcpyDeviceToHost));
        return AGX_OK;
    } catch(std::exception& e) { return fail(AGX_EDEVICE, e.what()); }
}

void agx_orbit_record_free(agx_orbit_record* R)
{
    if(R) (void)hipSetDevice(R->device);
    delete R;
}

}  // extern "C"

// ---------------- target plug-ins ----------------
struct agx_target {
    int type = 0, lmax = 0, mmax = 0;
    double axis[3] = {1, 1, 1};       // DensityClassic: shell axes
    std::vector<double> gridr;        // radial nodes without the origin
    std::vector<double> gridz;        // vertical nodes without z = 0 (cylindrical targets)
    int numValues = 0;
};

extern "C" {

agx_target* agx_target_density_sphharm(int lmax, int mmax, const double* gridr, int ngridr)
{
    g_err.clear();
    // TargetDensitySphHarm ctor (galaxymodel_densitygrid.cpp:356-369): a grid starting at zero
    // loses that node (the origin is always implied)
    if(!gridr || ngridr < 1) { fail(AGX_EINVAL, "TargetDensitySphHarm: invalid grid parameters"); return nullptr; }
    std::unique_ptr<agx_target> T(new agx_target);
    T->type = AGX_TARGET_DENSITY_SPHHARM;
    T->lmax = lmax; T->mmax = mmax;
    T->gridr.assign(gridr[0] > 0 ? gridr : gridr + 1, gridr + ngridr);
    bool ok = lmax >= 0 && mmax >= 0 && mmax <= lmax && T->gridr.size() >= 1 && T->gridr[0] > 0;
    for(size_t i = 1; i < T->gridr.size(); i++) ok &= T->gridr[i] > T->gridr[i-1];
    if(!ok) { fail(AGX_EINVAL, "TargetDensitySphHarm: invalid grid parameters"); return nullptr; }
    const int angular = (lmax/2+1) * (mmax/2+1) - mmax/2 * (mmax/2+1) / 2;
    T->numValues = angular * (int)T->gridr.size() + 1;
    return T.release();
}

agx_target* agx_target_density_cylindrical(int degree, int mmax, const double* gridR, int ngridR,
    const double* gridz, int ngridz)
{
    g_err.clear();
    // TargetDensityCylindrical<N> ctor (galaxymodel_densitygrid.cpp:486-503): grids starting at 0
    // lose that node
    if((degree != 0 && degree != 1) || !gridR || !gridz || ngridR < 1 || ngridz < 1) {
        fail(AGX_EINVAL, "TargetDensityCylindrical: invalid grid parameters");
        return nullptr;
    }
    std::unique_ptr<agx_target> T(new agx_target);
    T->type = degree == 0 ? AGX_TARGET_DENSITY_CYL_TOPHAT : AGX_TARGET_DENSITY_CYL_LINEAR;
    T->mmax = mmax;
    T->gridr.assign(gridR[0] > 0 ? gridR : gridR + 1, gridR + ngridR);
    T->gridz.assign(gridz[0] > 0 ? gridz : gridz + 1, gridz + ngridz);
    bool ok = mmax >= 0 && T->gridr.size() >= 1 && T->gridz.size() >= 1 && T->gridr[0] > 0 && T->gridz[0] > 0;
    for(size_t i = 1; i < T->gridr.size(); i++) ok &= T->gridr[i] > T->gridr[i-1];
    for(size_t i = 1; i < T->gridz.size(); i++) ok &= T->gridz[i] > T->gridz[i-1];
    if(!ok) { fail(AGX_EINVAL, "TargetDensityCylindrical: invalid grid parameters"); return nullptr; }
    T->numValues = ((int)T->gridz.size() + degree) * ((int)T->gridr.size() * (mmax/2+1) + degree);
    return T.release();
}

agx_target* agx_target_density_classic(int degree, int stripsPerPane, const double* gridr, int ngridr,
    double axisRatioY, double axisRatioZ)
{
    g_err.clear();
    // TargetDensityClassic<N> ctor (galaxymodel_densitygrid.cpp:148-167): a grid starting at 0
    // loses that node; axes scaled to a unit product
    if((degree != 0 && degree != 1) || !gridr || ngridr < 1) {
        fail(AGX_EINVAL, "TargetDensityClassic: invalid grid parameters");
        return nullptr;
    }
    std::unique_ptr<agx_target> T(new agx_target);
    T->type = degree == 0 ? AGX_TARGET_DENSITY_CLASSIC_TOPHAT : AGX_TARGET_DENSITY_CLASSIC_LINEAR;
    T->mmax = stripsPerPane;
    T->gridr.assign(gridr[0] > 0 ? gridr : gridr + 1, gridr + ngridr);
    bool ok = T->gridr.size() >= 1 && T->gridr[0] > 0 && stripsPerPane >= 1 && axisRatioY > 0 && axisRatioZ > 0;
    for(size_t i = 1; i < T->gridr.size(); i++) ok &= T->gridr[i] > T->gridr[i-1];
    if(!ok) { fail(AGX_EINVAL, "TargetDensityClassic: invalid grid parameters"); return nullptr; }
    T->axis[0] = 1. / cbrt(axisRatioY * axisRatioZ);
    T->axis[1] = axisRatioY * T->axis[0];
    T->axis[2] = axisRatioZ * T->axis[0];
    const int S = stripsPerPane;
    T->numValues = (3 * S * (S + degree) + degree) * (int)T->gridr.size() + degree;
    return T.release();
}

agx_target* agx_target_kinem_shell(int degree, const double* gridr, int ngridr)
{
    g_err.clear();
    // TargetKinemShell<N> (galaxymodel_losvd.h:121-136): B-splines of degree N on gridr as given
    if(degree < 0 || degree > 3 || !gridr || ngridr < 2) {
        fail(AGX_EINVAL, "KinemShell: degree must be 0..3 and gridr must have >= 2 nodes");
        return nullptr;
    }
    std::unique_ptr<agx_target> T(new agx_target);
    T->type = AGX_TARGET_KINEM_SHELL;
    T->lmax = degree;
    T->gridr.assign(gridr, gridr + ngridr);
    for(int i = 0; i < ngridr; i++)
        if(!(gridr[i] > -INFINITY && gridr[i] < INFINITY) || (i > 0 && !(gridr[i] > gridr[i-1]))) {
            fail(AGX_EINVAL, "BsplineInterpolator1d: x coordinates must be finite and monotonically increasing");
            return nullptr;
        }
    T->numValues = 2 * (ngridr + degree - 1);
    return T.release();
}

int agx_target_num_values(const agx_target* t) { return t ? t->numValues : -1; }

void agx_target_destroy(agx_target* t) { delete t; }

int agx_orbit_target_traj_batch(const agx_potential* pot, long n, const double* ic, const double* tstart,
    const double* ttotal, const int* trajsize, double accuracy, double Omega, long maxNumSteps, int ntargets,
    const agx_target* const* targets, float* const* out_targets, double* out_time, double* out_traj,
    double* out_final, long* out_nsteps, int* out_status, int flags);

int agx_orbit_target_batch(const agx_potential* pot, long n, const double* ic, const double* tstart,
    const double* ttotal, double accuracy, double Omega, long maxNumSteps, int ntargets,
    const agx_target* const* targets, float* const* out_targets, double* out_final, long* out_nsteps,
    int* out_status, int flags)
{
    return agx_orbit_target_traj_batch(pot, n, ic, tstart, ttotal, nullptr, accuracy, Omega, maxNumSteps, ntargets,
        targets, out_targets, nullptr, nullptr, out_final, out_nsteps, out_status, flags);
}

int agx_orbit_target_traj_batch(const agx_potential* pot, long n, const double* ic, const double* tstart,
    const double* ttotal, const int* trajsize, double accuracy, double Omega, long maxNumSteps, int ntargets,
    const agx_target* const* targets, float* const* out_targets, double* out_time, double* out_traj,
    double* out_final, long* out_nsteps, int* out_status, int flags)
{
    g_err.clear();
    if(!pot) return fail(AGX_EINVAL, "NULL potential");
    if(n < 0 || (n > 0 && (!ic || !ttotal))) return fail(AGX_EINVAL, "invalid orbit arrays");
    if(ntargets < 1 || ntargets > AGX_MAX_TARGETS || !targets || !out_targets)
        return fail(AGX_EINVAL, "between 1 and 4 targets with their outputs are required");
    for(int t = 0; t < ntargets; t++)
        if(!targets[t] || !out_targets[t]) return fail(AGX_EINVAL, "NULL target or output");
    if(trajsize)
        for(long i = 0; i < n; i++)
            if(trajsize[i] < 1) return fail(AGX_EINVAL, "trajsize must be >= 1 with targets");
    if(n == 0) return AGX_OK;
    try {
        HIPCHECK(hipSetDevice(pot->device));
        AgxTargetList TL;
        TL.count = ntargets;
        for(int t = 0; t < ntargets; t++) TL.needVelocity |= targets[t]->type == AGX_TARGET_KINEM_SHELL;
        std::vector<std::unique_ptr<DevBuf<double> > > grids, cubes;
        for(int t = 0; t < ntargets; t++) {
            const agx_target& T = *targets[t];
            grids.emplace_back(new DevBuf<double>(T.gridr.size() + T.gridz.size()));
            cubes.emplace_back(new DevBuf<double>((size_t)n * (T.numValues + 1)));
            HIPCHECK(hipMemcpy(grids.back()->p, T.gridr.data(), T.gridr.size() * 8, hipMemcpyHostToDevice));
            if(!T.gridz.empty())
                HIPCHECK(hipMemcpy(grids.back()->p + T.gridr.size(), T.gridz.data(), T.gridz.size() * 8,
                    hipMemcpyHostToDevice));
            HIPCHECK(hipMemset(cubes.back()->p, 0, (size_t)n * (T.numValues + 1) * 8));
            AgxTargetDev& D = TL.t[t];
            D.type = T.type; D.lmax = T.lmax; D.mmax = T.mmax; D.nr = (int)T.gridr.size();
            D.nz = (int)T.gridz.size();
            D.numValues = T.numValues; D.gridr = grids.back()->p; D.gridz = grids.back()->p + T.gridr.size();
            D.cube = cubes.back()->p;
            for(int c = 0; c < 3; c++) D.axis[c] = T.axis[c];
        }
        DevBuf<AgxTargetList> dTL(1);
        HIPCHECK(hipMemcpy(dTL.p, &TL, sizeof(TL), hipMemcpyHostToDevice));
        DevBuf<double> dic(n * 6), dts(tstart ? n : 0), dtt(n), dfin(out_final ? n * 6 : 0);
        DevBuf<int> dsz(n), dst(out_status ? n : 0);
        DevBuf<long> dns(out_nsteps ? n : 0);
        // trajectory rows as in agx_orbit_batch: trajsize[i] rows per orbit, in orbit order
        std::vector<int> sz(trajsize ? trajsize : nullptr, trajsize ? trajsize + n : nullptr);
        if(!trajsize) sz.assign(n, 1);
        std::vector<long> off(n);
        long rows = 0;
        for(long i = 0; i < n; i++) { off[i] = rows; rows += sz[i]; }
        const bool traj = trajsize && (out_time || out_traj);
        DevBuf<long> doff(traj ? n : 0);
        DevBuf<double> dtime(traj && out_time ? rows : 0), dtraj(traj && out_traj ? rows * 6 : 0);
        HIPCHECK(hipMemcpy(dic.p, ic, n * 48, hipMemcpyHostToDevice));
        if(tstart) HIPCHECK(hipMemcpy(dts.p, tstart, n * 8, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(dtt.p, ttotal, n * 8, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(dsz.p, sz.data(), n * 4, hipMemcpyHostToDevice));
        if(traj) HIPCHECK(hipMemcpy(doff.p, off.data(), n * 8, hipMemcpyHostToDevice));
        int rc = orbitBatchDevice(pot, n, dic.p, dts.p, dtt.p, dsz.p, traj ? doff.p : nullptr, nullptr, accuracy,
            Omega, maxNumSteps, dtime.p, dtraj.p, dfin.p, nullptr, dns.p, nullptr, nullptr, nullptr, dst.p, flags,
            nullptr, LyapDev(), RecordPool(), dTL.p);
        if(rc) return rc;
        HIPCHECK(hipDeviceSynchronize());
        // ~RuntimeFncTarget (galaxymodel_target.h:129-137): the datacube in storage precision,
        // times 1/time (itself rounded to storage precision); nothing when no time elapsed
        for(int t = 0; t < ntargets; t++) {
            const int nv = targets[t]->numValues;
            std::vector<double> cube((size_t)n * (nv + 1));
            HIPCHECK(hipMemcpy(cube.data(), cubes[t]->p, cube.size() * 8, hipMemcpyDeviceToHost));
            for(long i = 0; i < n; i++) {
                const double* row = &cube[(size_t)i * (nv + 1)];
                const double time = row[nv];
                float* out = out_targets[t] + (size_t)i * nv;
                for(int j = 0; j < nv; j++) out[j] = static_cast<float>(row[j]);
                if(time == 0) continue;
                const float invtime = static_cast<float>(1. / time);
                for(int j = 0; j < nv; j++) out[j] *= invtime;
            }
        }
        if(out_final) HIPCHECK(hipMemcpy(out_final, dfin.p, n * 48, hipMemcpyDeviceToHost));
        if(out_nsteps) HIPCHECK(hipMemcpy(out_nsteps, dns.p, n * 8, hipMemcpyDeviceToHost));
        if(out_status) HIPCHECK(hipMemcpy(out_status, dst.p, n * 4, hipMemcpyDeviceToHost));
        if(traj && out_time) HIPCHECK(hipMemcpy(out_time, dtime.p, rows * 8, hipMemcpyDeviceToHost));
        if(traj && out_traj) HIPCHECK(hipMemcpy(out_traj, dtraj.p, rows * 48, hipMemcpyDeviceToHost));
        return AGX_OK;
    } catch(std::exception& e) { return fail(AGX_EDEVICE, e.what()); }
}

}  // extern "C"

// ---------------- MLP surrogate (no counterpart in the reference) ----------------
struct agx_surrogate {
    const agx_potential* pot = nullptr;         // the full FP64 Multipole (not owned)
    std::unique_ptr<agx_potential> base;        // its l <= baseLmax truncation
    std::vector<float> theta;                   // flat parameters (AGX_P_* layout)
    int baseLmax = 2;
    float lnScale = 1, outScale = 1;
    double rmin = 1e-2, rmax = 1e2;
    std::vector<float> lossHist;
    void* dmem = nullptr;                       // packed inference weights + header
    AgxMlpDev* dM = nullptr;
    ~agx_surrogate() { if(dmem) (void)hipFree(dmem); }
};

namespace {

uint16_t bf16Bits(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7FFF + ((u >> 16) & 1);   // round to nearest even (finite inputs)
    return (uint16_t)(u >> 16);
}

// inference layout of the parameters (agx_mlp.h) uploaded into one allocation
void packInference(agx_surrogate& S)
{
    const float* th = S.theta.data();
    const size_t tileBytes = 8 * 16 * 64 * 16;    // one layer: 8 m-tiles x 16 k-blocks x 64 lanes x 16 B
    const size_t nf = 256 * 4 + 256 + 3 * 256 + 256;
    std::vector<uint16_t> tiles(6 * tileBytes / 2);
    const int woff[3] = { AGX_P_W2, AGX_P_W3, AGX_P_W4 }, boff[3] = { AGX_P_B2, AGX_P_B3, AGX_P_B4 };
    for(int l = 0; l < 3; l++) {
        const float* W = th + woff[l];
        uint16_t* F = tiles.data() + (size_t)l * tileBytes / 2;
        uint16_t* B = tiles.data() + (size_t)(3 + l) * tileBytes / 2;
        for(int mt = 0; mt < 8; mt++) for(int kb = 0; kb < 16; kb++) for(int lane = 0; lane < 64; lane++)
            for(int p = 0; p < 8; p++) {
                const size_t e = (((size_t)mt * 16 + kb) * 64 + lane) * 8 + p;
                const int row = 32 * mt + (lane & 31), k = agxMlpKNeuron(16 * kb + 8 * (lane >> 5) + p);
                F[e] = bf16Bits(W[row * 256 + k]);        // W[out = row][in = k]
                B[e] = bf16Bits(W[k * 256 + row]);        // W^T[in = row][out = k]
            }
    }
    std::vector<float> fl(nf);
    float* w1p = fl.data();
    float* b1p = w1p + 1024;
    float* bp = b1p + 256;
    float* w5p = bp + 768;
    for(int mt = 0; mt < 8; mt++) for(int g = 0; g < 2; g++) for(int j = 0; j < 16; j++) {
        const int idx = (mt * 2 + g) * 16 + j, o = agxMlpNeuron(mt, g, j);
        for(int f = 0; f < 4; f++) w1p[idx * 4 + f] = th[AGX_P_W1 + o * 4 + f];
        b1p[idx] = th[AGX_P_B1 + o];
        for(int l = 0; l < 3; l++) bp[l * 256 + idx] = th[boff[l] + o];
        w5p[idx] = th[AGX_P_W5 + o];
    }
    const size_t oT = 0, oF = oT + 6 * tileBytes, oH = oF + align256(nf * 4), total = oH + align256(sizeof(AgxMlpDev));
    HIPCHECK(hipSetDevice(S.pot->device));
    if(!S.dmem) HIPCHECK(hipMalloc(&S.dmem, total));
    char* base = static_cast<char*>(S.dmem);
    HIPCHECK(hipMemcpy(base + oT, tiles.data(), 6 * tileBytes, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(base + oF, fl.data(), nf * 4, hipMemcpyHostToDevice));
    AgxMlpDev M;
    memset(&M, 0, sizeof(M));
    for(int l = 0; l < 3; l++) {
        M.fwdA[l] = reinterpret_cast<const agx_bf16x8*>(base + oT + l * tileBytes);
        M.bwdA[l] = reinterpret_cast<const agx_bf16x8*>(base + oT + (3 + l) * tileBytes);
        M.bp[l] = reinterpret_cast<const float*>(base + oF) + 1024 + 256 + l * 256;
    }
    M.w1p = reinterpret_cast<const float*>(base + oF);
    M.b1p = M.w1p + 1024;
    M.w5p = M.w1p + 1024 + 256 + 768;
    M.b5 = th[AGX_P_B5];
    M.lnScale = S.lnScale;
    M.outScale = S.outScale;
    M.rminTrain = S.rmin;
    M.rmaxTrain = S.rmax;
    S.dM = reinterpret_cast<AgxMlpDev*>(base + oH);
    HIPCHECK(hipMemcpy(S.dM, &M, sizeof(M), hipMemcpyHostToDevice));
}

agx_surrogate* makeSurrogate(const agx_potential* pot, int baseLmax, float lnScale, float outScale,
    double rmin, double rmax)
{
    if(baseLmax < 0 || baseLmax > 8) throw std::invalid_argument("surrogate: base_lmax must be 0 to 8");
    if(!(rmin > 0 && rmax > rmin)) throw std::invalid_argument("surrogate: invalid radius range");
    HIPCHECK(hipSetDevice(pot->device));
    std::unique_ptr<agx_surrogate> S(new agx_surrogate);
    S->pot = pot;
    S->baseLmax = baseLmax;
    S->lnScale = lnScale;
    S->outScale = outScale;
    S->rmin = rmin;
    S->rmax = rmax;
    S->base.reset(makeFromCoefs(agx::truncateCoefs(pot->coefs, baseLmax), true, pot->nfwMass, pot->nfwScale));
    return S.release();
}

// device buffers of the training step (freed on destruction)
struct TrainAlloc {
    std::vector<void*> ptrs;
    template<typename T> T* get(size_t n) {
        void* p = nullptr;
        HIPCHECK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
        ptrs.push_back(p);
        return static_cast<T*>(p);
    }
    ~TrainAlloc() { for(void* p : ptrs) (void)hipFree(p); }
};

void allocTrainStep(TrainAlloc& A, AgxTrainBuffers& T, int batch, long nbatches, long histLen)
{
    const size_t E = (size_t)batch * 256;
    T.batch = batch;
    T.nbatches = nbatches;
    T.theta = A.get<float>(AGX_MLP_NPARAM);
    T.grad = A.get<float>(AGX_MLP_NPARAM);
    T.m1 = A.get<float>(AGX_MLP_NPARAM);
    T.m2 = A.get<float>(AGX_MLP_NPARAM);
    T.Wb = A.get<uint16_t>(3 * 65536);
    for(int l = 0; l < 4; l++) { T.H[l] = A.get<uint16_t>(E); T.S[l] = A.get<float>(E); }
    T.Z = A.get<float>(E);
    T.G = A.get<float>(E);
    T.Gb = A.get<uint16_t>(E);
    T.dout = A.get<float>(batch);
    T.step = A.get<long>(1);
    T.lossHist = A.get<float>(histLen);
    HIPCHECK(hipMemset(T.m1, 0, AGX_MLP_NPARAM * 4));
    HIPCHECK(hipMemset(T.m2, 0, AGX_MLP_NPARAM * 4));
    HIPCHECK(hipMemset(T.step, 0, 8));
    HIPCHECK(hipMemset(T.lossHist, 0, histLen * 4));
}

std::vector<float> initParams(unsigned long long seed)
{
    std::vector<float> th(AGX_MLP_NPARAM, 0.f);
    std::mt19937_64 gen(seed);
    std::normal_distribution<double> nd(0., 1.);
    for(int i = 0; i < 256 * 4; i++) th[AGX_P_W1 + i] = (float)(nd(gen) * sqrt(2. / 4));
    const int woff[3] = { AGX_P_W2, AGX_P_W3, AGX_P_W4 };
    for(int l = 0; l < 3; l++) for(int i = 0; i < 65536; i++) th[woff[l] + i] = (float)(nd(gen) * sqrt(2. / 256));
    for(int i = 0; i < 256; i++) th[AGX_P_W5 + i] = (float)(nd(gen) * sqrt(1. / 256));
    return th;
}

}  // namespace

extern "C" {

agx_surrogate* agx_surrogate_create(const agx_potential* pot, const float* theta, int base_lmax,
    float ln_scale, float out_scale, double rmin, double rmax)
{
    g_err.clear();
    if(!pot || !theta) { fail(AGX_EINVAL, "NULL argument"); return nullptr; }
    try {
        std::unique_ptr<agx_surrogate> S(makeSurrogate(pot, base_lmax, ln_scale, out_scale, rmin, rmax));
        S->theta.assign(theta, theta + AGX_MLP_NPARAM);
        packInference(*S);
        return S.release();
    } catch(std::exception& e) { fail(AGX_EINVAL, e.what()); return nullptr; }
}

agx_surrogate* agx_surrogate_train(const agx_potential* pot, const char* params)
{
    g_err.clear();
    if(!pot) { fail(AGX_EINVAL, "NULL potential"); return nullptr; }
    try {
        agx::Params P = agx::Params::parse(params ? params : "");
        P.checkOnly({"samples", "batch", "steps", "lr", "seed", "baselmax", "rmin", "rmax", "warmup"});
        const int batch = (int)P.num("batch", 16384);
        const long steps = (long)P.num("steps", 6000);
        const double rmin = P.num("rmin", 1e-2), rmax = P.num("rmax", 1e2);
        const unsigned long long seed = (unsigned long long)P.num("seed", 1);
        if(batch < 512 || batch % 512) throw std::invalid_argument("surrogate: batch must be a multiple of 512");
        if(steps < 1) throw std::invalid_argument("surrogate: steps must be positive");
        const long nb = std::max<long>(1, (long)P.num("samples", 20e6) / batch), N = nb * batch;
        const float lnScale = (float)std::max(fabs(log(rmin)), fabs(log(rmax)));
        std::unique_ptr<agx_surrogate> S(makeSurrogate(pot, (int)P.num("baselmax", 2), lnScale, 1.f, rmin, rmax));
        hipStream_t st;
        HIPCHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        std::unique_ptr<void, void(*)(void*)> stGuard((void*)st, [](void* s) { (void)hipStreamDestroy((hipStream_t)s); });
        TrainAlloc A;
        float* feat = A.get<float>(N * 4);
        float* target = A.get<float>(N);
        {   // training set: FP64 Multipole and its l <= baseLmax part at N points
            TrainAlloc tmp;
            double* pos = tmp.get<double>(N * 3);
            double* phi = tmp.get<double>(N);
            double* phib = tmp.get<double>(N);
            double* q = tmp.get<double>(N);
            double* sumsq = tmp.get<double>(1);
            HIPCHECK(hipMemsetAsync(sumsq, 0, 8, st));
            HIPCHECK(agxLaunchSamplePoints(N, seed, rmin, rmax, pos, st));
            HIPCHECK(agxLaunchEval(pot->dev, pot->ddev, AGX_FP64, AGX_FETCH_VEC2, N, pos, phi, nullptr, st));
            HIPCHECK(agxLaunchEval(S->base->dev, S->base->ddev, AGX_FP64, AGX_FETCH_VEC2, N, pos, phib, nullptr, st));
            HIPCHECK(agxLaunchResidual(N, pos, phi, phib, lnScale, feat, q, sumsq, st));
            double ss = 0;
            HIPCHECK(hipMemcpyAsync(&ss, sumsq, 8, hipMemcpyDeviceToHost, st));
            HIPCHECK(hipStreamSynchronize(st));
            const double rms = sqrt(ss / N);
            if(!(rms > 0)) throw std::runtime_error("surrogate: the potential has no l > base_lmax part to learn");
            S->outScale = (float)rms;
            HIPCHECK(agxLaunchTarget(N, q, 1 / (double)S->outScale, target, st));
            HIPCHECK(hipStreamSynchronize(st));
        }
        AgxTrainBuffers T;
        allocTrainStep(A, T, batch, nb, steps);
        T.feat = feat;
        T.target = target;
        S->theta = initParams(seed);
        HIPCHECK(hipMemcpy(T.theta, S->theta.data(), AGX_MLP_NPARAM * 4, hipMemcpyHostToDevice));
        HIPCHECK(agxLaunchPackTrain(T.theta, T.Wb, st));
        AgxTrainHyper hp;
        hp.lr = (float)P.num("lr", 1e-3);
        hp.beta1 = 0.9f; hp.beta2 = 0.999f; hp.eps = 1e-8f;
        hp.totalSteps = steps;
        hp.warmup = (long)P.num("warmup", 200);
        // one training step captured as a graph and replayed
        hipGraph_t graph;
        hipGraphExec_t exec;
        HIPCHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        hipError_t ce = agxEnqueueTrainStep(T, hp, st);
        HIPCHECK(hipStreamEndCapture(st, &graph));
        HIPCHECK(ce);
        std::unique_ptr<void, void(*)(void*)> gGuard((void*)graph, [](void* g) { (void)hipGraphDestroy((hipGraph_t)g); });
        HIPCHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        std::unique_ptr<void, void(*)(void*)> eGuard((void*)exec, [](void* g) { (void)hipGraphExecDestroy((hipGraphExec_t)g); });
        for(long s = 0; s < steps; s++) HIPCHECK(hipGraphLaunch(exec, st));
        HIPCHECK(hipStreamSynchronize(st));
        HIPCHECK(hipMemcpy(S->theta.data(), T.theta, AGX_MLP_NPARAM * 4, hipMemcpyDeviceToHost));
        S->lossHist.resize(steps);
        HIPCHECK(hipMemcpy(S->lossHist.data(), T.lossHist, steps * 4, hipMemcpyDeviceToHost));
        packInference(*S);
        return S.release();
    } catch(std::exception& e) { fail(AGX_EDEVICE, e.what()); return nullptr; }
}

void agx_surrogate_destroy(agx_surrogate* s) { delete s; }

int agx_surrogate_info(const agx_surrogate* s, float* theta, double* meta)
{
    if(!s) return fail(AGX_EINVAL, "NULL surrogate");
    if(theta) memcpy(theta, s->theta.data(), AGX_MLP_NPARAM * 4);
    if(meta) {
        meta[0] = s->baseLmax; meta[1] = s->lnScale; meta[2] = s->outScale; meta[3] = s->rmin;
        meta[4] = s->rmax; meta[5] = (double)s->lossHist.size();
        meta[6] = s->lossHist.empty() ? NAN : s->lossHist.back();
        meta[7] = AGX_MLP_NPARAM;
    }
    return AGX_OK;
}

long agx_surrogate_loss_history(const agx_surrogate* s, float* out, long cap)
{
    if(!s) return -fail(AGX_EINVAL, "NULL surrogate");
    const long n = std::min<long>(cap, (long)s->lossHist.size());
    if(out && n > 0) memcpy(out, s->lossHist.data(), n * 4);
    return (long)s->lossHist.size();
}

int agx_surrogate_eval_many_device(const agx_surrogate* s, long n, const double* d_pos, double* d_phi,
    double* d_grad, void* stream)
{
    if(!s) return fail(AGX_EINVAL, "NULL surrogate");
    if(n < 0 || (n > 0 && !d_pos)) return fail(AGX_EINVAL, "invalid arrays");
    hipError_t e = agxLaunchSurrogateEval(s->dM, s->base->ddev, n, d_pos, d_phi, d_grad, (hipStream_t)stream);
    return e == hipSuccess ? AGX_OK : fail(AGX_EDEVICE, hipGetErrorString(e));
}

int agx_surrogate_eval_many(const agx_surrogate* s, long n, const double* pos, double* phi, double* grad)
{
    g_err.clear();
    if(!s) return fail(AGX_EINVAL, "NULL surrogate");
    if(n == 0) return AGX_OK;
    if(n < 0 || !pos) return fail(AGX_EINVAL, "invalid arrays");
    try {
        HIPCHECK(hipSetDevice(s->pot->device));
        DevBuf<double> dp(n * 3), dphi(phi ? n : 0), dg(grad ? n * 3 : 0);
        HIPCHECK(hipMemcpy(dp.p, pos, n * 24, hipMemcpyHostToDevice));
        HIPCHECK(agxLaunchSurrogateEval(s->dM, s->base->ddev, n, dp.p, dphi.p, dg.p, nullptr));
        HIPCHECK(hipDeviceSynchronize());
        if(phi) HIPCHECK(hipMemcpy(phi, dphi.p, n * 8, hipMemcpyDeviceToHost));
        if(grad) HIPCHECK(hipMemcpy(grad, dg.p, n * 24, hipMemcpyDeviceToHost));
        return AGX_OK;
    } catch(std::exception& e) { return fail(AGX_EDEVICE, e.what()); }
}

int agx_surrogate_network(const agx_surrogate* s, long n, const float* feat, float* out, float* dfeat)
{
    g_err.clear();
    if(!s) return fail(AGX_EINVAL, "NULL surrogate");
    if(n == 0) return AGX_OK;
    if(n < 0 || !feat) return fail(AGX_EINVAL, "invalid arrays");
    try {
        HIPCHECK(hipSetDevice(s->pot->device));
        DevBuf<float> df(n * 4), dout(out ? n : 0), dd(dfeat ? n * 4 : 0);
        HIPCHECK(hipMemcpy(df.p, feat, n * 16, hipMemcpyHostToDevice));
        HIPCHECK(agxLaunchMlpFeatures(s->dM, n, df.p, dout.p, dd.p, nullptr));
        HIPCHECK(hipDeviceSynchronize());
        if(out) HIPCHECK(hipMemcpy(out, dout.p, n * 4, hipMemcpyDeviceToHost));
        if(dfeat) HIPCHECK(hipMemcpy(dfeat, dd.p, n * 16, hipMemcpyDeviceToHost));
        return AGX_OK;
    } catch(std::exception& e) { return fail(AGX_EDEVICE, e.what()); }
}

int agx_surrogate_train_grad(const agx_surrogate* s, long n, const float* feat, const float* target,
    float* loss, float* grad)
{
    g_err.clear();
    if(!s) return fail(AGX_EINVAL, "NULL surrogate");
    if(n < 512 || n % 512 || !feat || !target) return fail(AGX_EINVAL, "n must be a positive multiple of 512");
    try {
        HIPCHECK(hipSetDevice(s->pot->device));
        TrainAlloc A;
        AgxTrainBuffers T;
        allocTrainStep(A, T, (int)n, 1, 1);
        float* df = A.get<float>(n * 4);
        float* dt = A.get<float>(n);
        HIPCHECK(hipMemcpy(df, feat, n * 16, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(dt, target, n * 4, hipMemcpyHostToDevice));
        T.feat = df;
        T.target = dt;
        HIPCHECK(hipMemcpy(T.theta, s->theta.data(), AGX_MLP_NPARAM * 4, hipMemcpyHostToDevice));
        HIPCHECK(agxLaunchPackTrain(T.theta, T.Wb, nullptr));
        AgxTrainHyper hp;
        memset(&hp, 0, sizeof(hp));   // lr = 0: loss and gradient only
        HIPCHECK(agxEnqueueTrainStep(T, hp, nullptr));
        HIPCHECK(hipDeviceSynchronize());
        if(loss) HIPCHECK(hipMemcpy(loss, T.lossHist, 4, hipMemcpyDeviceToHost));
        if(grad) HIPCHECK(hipMemcpy(grad, T.grad, AGX_MLP_NPARAM * 4, hipMemcpyDeviceToHost));
        return AGX_OK;
    } catch(std::exception& e) { return fail(AGX_EDEVICE, e.what()); }
}

int agx_surrogate_orbit_batch_device(const agx_surrogate* s, long n, const double* d_ic,
    const double* d_tstart, const double* d_ttotal, const int* d_trajsize, const long* d_traj_offset,
    double accuracy, double Omega, long maxNumSteps, double max_rel_de, double* d_time, double* d_traj,
    double* d_final, long* d_nsteps, long* d_nattempts, long* d_nevals, double* d_maxde, int* d_status,
    long* n_fallback, void* stream)
{
    if(!s) return fail(AGX_EINVAL, "NULL surrogate");
    if(n_fallback) *n_fallback = 0;
    if(n == 0) return AGX_OK;
    if(n < 0 || n > 0x7fffffffL || !d_ic || !d_ttotal || !d_trajsize)
        return fail(AGX_EINVAL, "invalid orbit arrays");
    if((d_time || d_traj) && !d_traj_offset) return fail(AGX_EINVAL, "trajectory output needs offsets");
    if(!(accuracy > 0)) return fail(AGX_EINVAL, "accuracy must be positive");
    if(!(Omega > -INFINITY && Omega < INFINITY)) return fail(AGX_EINVAL, "Omega must be finite");
    if(maxNumSteps <= 0) return fail(AGX_EINVAL, "maxNumSteps must be positive");
    if(!(max_rel_de > 0)) return fail(AGX_EINVAL, "max_rel_de must be positive");
    const agx_potential* pot = s->pot;
    try {
        hipStream_t st = (hipStream_t)stream;
        const size_t schedBytes = agxScheduleWorkspace(n), ordBytes = ((size_t)n * 4 + 255) / 256 * 256;
        void* work = nullptr;
        HIPCHECK(hipMallocAsync(&work, schedBytes + 2 * ordBytes + 256, st));
        char* w = static_cast<char*>(work);
        int* ord = reinterpret_cast<int*>(w + schedBytes);
        int* list = reinterpret_cast<int*>(w + schedBytes + ordBytes);
        unsigned int* nfb = reinterpret_cast<unsigned int*>(w + schedBytes + 2 * ordBytes);
        auto release = [&]() { (void)hipFreeAsync(work, st); };
        hipError_t e = agxLaunchSchedule(pot->dev, pot->ddev, n, d_ic, d_ttotal, ord, work, st);
        if(e != hipSuccess) { release(); return fail(AGX_EDEVICE, std::string("schedule: ") + hipGetErrorString(e)); }
        AgxOrbitArgs A;
        A.n = n; A.ic = d_ic; A.tstart = d_tstart; A.ttotal = d_ttotal; A.trajsize = d_trajsize;
        A.trajOffset = d_traj_offset; A.order = ord; A.accuracy = accuracy; A.Omega = Omega;
        A.maxNumSteps = maxNumSteps; A.outTime = d_time; A.outTraj = d_traj; A.outFinal = d_final;
        A.outNsteps = d_nsteps; A.outNattempts = d_nattempts; A.outNevals = d_nevals;
        A.outMaxde = d_maxde; A.outStatus = d_status; A.counter = pot->dcounter;
        AgxSurrogateOrbitArgs SA;
        SA.maxRelDE = max_rel_de; SA.rmin = s->rmin; SA.rmax = s->rmax; SA.fallback = list; SA.nfallback = nfb;
        int dev = 0, cus = 0;
        HIPCHECK(hipGetDevice(&dev));
        HIPCHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        const long need = (n + 63) / 64;
        const int blocks = (int)std::min<long>(need, (long)cus * agxOrbitSurrogateBlocksPerCU());
        // Pilot (batches of 8192 orbits or more): every 64th orbit for at most 200 steps on the
        // surrogate, outputs discarded.  When half of them or more already exceed max_rel_de or
        // leave the training radii, the surrogate cannot hold this batch to the tolerance and
        // every orbit goes to the FP64 Multipole directly (all marked as fallback) instead of
        // being integrated twice.
        bool routeAll = false;
        g_lastPilot[0] = g_lastPilot[1] = 0;
        if(n >= 8192 && g_surrogatePilot) {
            const long np = std::max(1024L, n / 64), stride = n / np;
            std::vector<int> pl(np);
            for(long j = 0; j < np; j++) pl[j] = (int)(j * stride);
            void* pw = nullptr;
            HIPCHECK(hipMallocAsync(&pw, (size_t)np * 4, st));
            HIPCHECK(hipMemcpyAsync(pw, pl.data(), (size_t)np * 4, hipMemcpyHostToDevice, st));
            AgxOrbitArgs PA = A;
            PA.n = np; PA.order = static_cast<int*>(pw);
            PA.maxNumSteps = std::min(maxNumSteps, 200L);
            PA.outTime = nullptr; PA.outTraj = nullptr; PA.outFinal = nullptr; PA.outNsteps = nullptr;
            PA.outNattempts = nullptr; PA.outNevals = nullptr; PA.outMaxde = nullptr; PA.outStatus = nullptr;
            const int pblocks = (int)std::min<long>((np + 63) / 64, (long)cus * agxOrbitSurrogateBlocksPerCU());
            HIPCHECK(agxLaunchFillU32(pot->dcounter, 1, 0u, st));
            HIPCHECK(agxLaunchFillU32(nfb, 1, 0u, st));
            AgxSurrogateOrbitArgs SP = SA;
            SP.abortAt = (unsigned int)((np + 1) / 2);   // the routing decision below is then known
            e = agxLaunchOrbitSurrogate(pot->ddev, s->base->ddev, s->dM, PA, SP, pblocks, st);
            unsigned int pf = 0;
            if(e == hipSuccess) e = hipMemcpyAsync(&pf, nfb, sizeof(pf), hipMemcpyDeviceToHost, st);
            if(e == hipSuccess) e = hipStreamSynchronize(st);
            (void)hipFreeAsync(pw, st);
            if(e != hipSuccess) { release(); return fail(AGX_EDEVICE, std::string("surrogate pilot: ") + hipGetErrorString(e)); }
            g_lastPilot[0] = np; g_lastPilot[1] = pf;   // pf stops growing at the decision
            routeAll = 2 * (long)pf >= np;
        }
        unsigned int nf = 0;
        if(routeAll) {
            nf = (unsigned int)n;
            list = ord;   // every orbit, longest first
        } else {
            HIPCHECK(agxLaunchFillU32(pot->dcounter, 1, 0u, st));
            HIPCHECK(agxLaunchFillU32(nfb, 1, 0u, st));
            e = agxLaunchOrbitSurrogate(pot->ddev, s->base->ddev, s->dM, A, SA, blocks, st);
            if(e != hipSuccess) { release(); return fail(AGX_EDEVICE, std::string("surrogate orbit launch: ") + hipGetErrorString(e)); }
            HIPCHECK(hipMemcpyAsync(&nf, nfb, sizeof(nf), hipMemcpyDeviceToHost, st));
            HIPCHECK(hipStreamSynchronize(st));
        }
        if(nf > 0) {   // the fallback orbits, integrated again with the FP64 Multipole
            // the FP64 batch path (work order, probe and group mode as for agx_orbit_batch_device):
            // the whole batch in its own order, or the fallback list in the order it was filled
            const int rc = orbitBatchDevice(pot, routeAll ? n : (long)nf, d_ic, d_tstart, d_ttotal, d_trajsize,
                d_traj_offset, routeAll ? nullptr : list, accuracy, Omega, maxNumSteps, d_time, d_traj, d_final,
                nullptr, d_nsteps, d_nattempts, d_nevals, d_maxde, d_status, AGX_PRECISION_FP64, st);
            if(rc != AGX_OK) { const std::string msg = g_err; release(); return fail(rc, "fallback: " + msg); }
            e = agxLaunchMarkFallback(list, nf, d_status, st);
            if(e != hipSuccess) { release(); return fail(AGX_EDEVICE, std::string("fallback launch: ") + hipGetErrorString(e)); }
        }
        release();
        if(n_fallback) *n_fallback = nf;
        return AGX_OK;
    } catch(std::exception& e) { return fail(AGX_EDEVICE, e.what()); }
}

int agx_set_surrogate_pilot(int enable)
{
    g_surrogatePilot = enable != 0;
    return AGX_OK;
}

int agx_surrogate_last_pilot(long* out_pilot_orbits, long* out_pilot_fallbacks)
{
    if(out_pilot_orbits) *out_pilot_orbits = g_lastPilot[0];
    if(out_pilot_fallbacks) *out_pilot_fallbacks = g_lastPilot[1];
    return AGX_OK;
}

int agx_surrogate_orbit_batch(const agx_surrogate* s, long n, const double* ic, const double* tstart,
    const double* ttotal, const int* trajsize, double accuracy, double Omega, long maxNumSteps,
    double max_rel_de, double* out_time, double* out_traj, double* out_final, long* out_nsteps,
    long* out_nattempts, long* out_nevals, double* out_maxde, int* out_status, long* n_fallback)
{
    g_err.clear();
    if(!s) return fail(AGX_EINVAL, "NULL surrogate");
    if(n_fallback) *n_fallback = 0;
    if(n == 0) return AGX_OK;
    if(n < 0 || !ic || !ttotal || !trajsize) return fail(AGX_EINVAL, "invalid orbit arrays");
    std::vector<long> offs(n);
    long rows = 0;
    for(long i = 0; i < n; i++) {
        if(trajsize[i] < 1) return fail(AGX_EUNSUPPORTED, "trajsize must be >= 1");
        offs[i] = rows;
        rows += trajsize[i];
    }
    try {
        HIPCHECK(hipSetDevice(s->pot->device));
        DevBuf<double> dic(n * 6), dts(tstart ? n : 0), dtt(n), dtime(out_time ? rows : 0),
            dtraj(out_traj ? rows * 6 : 0), dfin(out_final ? n * 6 : 0), dmde(out_maxde ? n : 0);
        DevBuf<int> dsz(n), dst(out_status ? n : 0);
        DevBuf<long> doff(n), dns(out_nsteps ? n : 0), dna(out_nattempts ? n : 0), dne(out_nevals ? n : 0);
        HIPCHECK(hipMemcpy(dic.p, ic, n * 48, hipMemcpyHostToDevice));
        if(tstart) HIPCHECK(hipMemcpy(dts.p, tstart, n * 8, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(dtt.p, ttotal, n * 8, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(dsz.p, trajsize, n * 4, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(doff.p, offs.data(), n * 8, hipMemcpyHostToDevice));
        int rc = agx_surrogate_orbit_batch_device(s, n, dic.p, dts.p, dtt.p, dsz.p, doff.p, accuracy, Omega,
            maxNumSteps, max_rel_de, dtime.p, dtraj.p, dfin.p, dns.p, dna.p, dne.p, dmde.p, dst.p, n_fallback, nullptr);
        if(rc) return rc;
        HIPCHECK(hipDeviceSynchronize());
        if(out_time) HIPCHECK(hipMemcpy(out_time, dtime.p, rows * 8, hipMemcpyDeviceToHost));
        if(out_traj) HIPCHECK(hipMemcpy(out_traj, dtraj.p, rows * 48, hipMemcpyDeviceToHost));
        if(out_final) HIPCHECK(hipMemcpy(out_final, dfin.p, n * 48, hipMemcpyDeviceToHost));
        if(out_nsteps) HIPCHECK(hipMemcpy(out_nsteps, dns.p, n * 8, hipMemcpyDeviceToHost));
        if(out_nattempts) HIPCHECK(hipMemcpy(out_nattempts, dna.p, n * 8, hipMemcpyDeviceToHost));
        if(out_nevals) HIPCHECK(hipMemcpy(out_nevals, dne.p, n * 8, hipMemcpyDeviceToHost));
        if(out_maxde) HIPCHECK(hipMemcpy(out_maxde, dmde.p, n * 8, hipMemcpyDeviceToHost));
        if(out_status) HIPCHECK(hipMemcpy(out_status, dst.p, n * 4, hipMemcpyDeviceToHost));
        return AGX_OK;
    } catch(std::exception& e) { return fail(AGX_EDEVICE, e.what()); }
}

int agx_debug_gemm(int M, int N, int K, int splits, int a_kcontig, int b_kcontig, const unsigned short* A,
    const unsigned short* B, float* C)
{
    g_err.clear();
    if(M <= 0 || N <= 0 || K <= 0 || !A || !B || !C) return fail(AGX_EINVAL, "invalid arguments");
    try {
        DevBuf<uint16_t> da((size_t)M * K), db((size_t)K * N);
        DevBuf<float> dc((size_t)M * N);
        HIPCHECK(hipMemcpy(da.p, A, (size_t)M * K * 2, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(db.p, B, (size_t)K * N * 2, hipMemcpyHostToDevice));
        HIPCHECK(hipMemset(dc.p, 0, (size_t)M * N * 4));
        hipError_t e = agxLaunchGemm(M, N, K, splits, a_kcontig != 0, b_kcontig != 0, da.p,
            a_kcontig ? K : M, db.p, b_kcontig ? K : N, dc.p, N, splits > 1 ? 2 : 0, nullptr);
        if(e != hipSuccess) return fail(AGX_EINVAL, std::string("gemm: ") + hipGetErrorString(e));
        HIPCHECK(hipDeviceSynchronize());
        HIPCHECK(hipMemcpy(C, dc.p, (size_t)M * N * 4, hipMemcpyDeviceToHost));
        return AGX_OK;
    } catch(std::exception& e) { return fail(AGX_EDEVICE, e.what()); }
}

// ---- drop-in aliases (src/interface_c.cpp:15-98) ----
const char* agama_getError(void) { return agx_last_error(); }
agama_Potential* agama_createPotential(const char* params) { return agx_potential_create(params); }
void agama_deletePotential(agama_Potential* pot) { agx_potential_destroy(pot); }
double agama_evalPotential(const agama_Potential* pot, const double pos[3], double time,
    double deriv[3], double deriv2[6])
{ return agx_eval_potential(pot, pos, time, deriv, deriv2); }

}  // extern "C"

// ---- NEMO acceleration plugin (src/interface_nemo.cpp:47-220) ----
namespace {
std::unique_ptr<agx_potential> g_nemoPot;
double g_nemoOmega = 0;

// getAcc (interface_nemo.cpp:74-97): the active bodies' positions rotated into the frame of the
// potential at angle Omega*time (the reference's math::sincos), one batched evaluation on the
// device, accelerations rotated back; potential (flag bit 0) and acceleration (bit 1) added
// or assigned
template<typename NumT>
void nemoAcc(int ndim, double time, int nbody, const void* pos_, const int* active, void* pot_,
    void* acc_, int flag)
{
    const NumT* pos = static_cast<const NumT*>(pos_);
    NumT* pot = static_cast<NumT*>(pot_);
    NumT* acc = static_cast<NumT*>(acc_);
    double ca = 1, sa = 0;
    agx::sincosAgama(g_nemoOmega * time, sa, ca);
    std::vector<long> idx;
    std::vector<double> pts;
    idx.reserve(nbody);
    pts.reserve((size_t)nbody * 3);
    for(int i = 0; i < nbody; i++) {
        if(active && (active[i] & 1) != 1) continue;
        idx.push_back(i);
        const double x = pos[i * ndim], y = pos[i * ndim + 1], z = ndim == 3 ? pos[i * ndim + 2] : 0;
        pts.push_back(x * ca + y * sa);
        pts.push_back(y * ca - x * sa);
        pts.push_back(z);
    }
    if(idx.empty()) return;
    std::vector<double> phi(idx.size()), g(idx.size() * 3);
    if(agx_eval_many(g_nemoPot.get(), (long)idx.size(), pts.data(), time, phi.data(), g.data(),
        AGX_PRECISION_FP64) != AGX_OK) {
        fprintf(stderr, "agama_amd NEMO plugin: %s\n", agx_last_error());
        return;
    }
    const bool addPot = flag & 1, addAcc = flag & 2;
    for(size_t j = 0; j < idx.size(); j++) {
        const long i = idx[j];
        const NumT ax = -(NumT)(g[j * 3] * ca - g[j * 3 + 1] * sa),
                   ay = -(NumT)(g[j * 3 + 1] * ca + g[j * 3] * sa), az = -(NumT)g[j * 3 + 2];
        if(pot) { if(addPot) pot[i] += (NumT)phi[j]; else pot[i] = (NumT)phi[j]; }
        if(addAcc) { acc[i * ndim] += ax; acc[i * ndim + 1] += ay; if(ndim == 3) acc[i * ndim + 2] += az; }
        else { acc[i * ndim] = ax; acc[i * ndim + 1] = ay; if(ndim == 3) acc[i * ndim + 2] = az; }
    }
}

// myacc (interface_nemo.cpp:114-147)
void nemoAccel(int ndim, double t, int n, const void*, const void* pos, const void*,
    const int* flags, void* pot, void* acc, int flag, char numtype)
{
    if(!g_nemoPot || (ndim != 2 && ndim != 3)) return;
    if(numtype == 'f') nemoAcc<float>(ndim, t, n, pos, flags, pot, acc, flag);
    else if(numtype == 'd') nemoAcc<double>(ndim, t, n, pos, flags, pot, acc, flag);
}
}  // namespace

// inipotential (interface_nemo.cpp:180-206): pars = {Omega, G}; the potential comes from a
// Multipole coefficient file.  G != 1 rescales masses in the reference's unit conversion, which
// this library does not provide: it stops with a message, as for a missing file.
extern "C" void inipotential(const int* npar, const double* pars, const char* file)
{
    if(!file || !file[0]) {
        fprintf(stderr, "agama_amd NEMO plugin: should provide the coefficient file in accfile=...\n");
        exit(1);
    }
    const int np = npar ? *npar : 0;
    g_nemoOmega = np >= 1 ? pars[0] : 0;
    const double G = np >= 2 ? pars[1] : 1;
    if(G != 1) {
        fprintf(stderr, "agama_amd NEMO plugin: G != 1 (unit conversion) is not supported\n");
        exit(1);
    }
    agx_potential* p = agx_potential_create((std::string("type=Multipole, file=") + file).c_str());
    if(!p) {
        fprintf(stderr, "agama_amd NEMO plugin: %s\n", agx_last_error());
        exit(1);
    }
    g_nemoPot.reset(p);
    fprintf(stderr, "agama_amd NEMO plugin: created a Multipole potential from %s", file);
    if(g_nemoOmega != 0) fprintf(stderr, " with pattern speed %g", g_nemoOmega);
    fprintf(stderr, "\n");
}

// iniacceleration (interface_nemo.cpp:209-220)
extern "C" void iniacceleration(const double* pars, int npar, const char* file,
    agx_acc_pter* accel, bool* need_m, bool* need_v)
{
    inipotential(&npar, pars, file);
    *accel = nemoAccel;
    if(need_m) *need_m = false;
    if(need_v) *need_v = false;
}

// potential_double / potential_float (interface_nemo.cpp:150-177): one body, assigned
extern "C" void potential_double(const int* ndim, const double* pos, double* acc, double* pot, const double* time)
{
    if(g_nemoPot && (*ndim == 2 || *ndim == 3)) nemoAcc<double>(*ndim, *time, 1, pos, nullptr, pot, acc, 0);
}

extern "C" void potential_float(const int* ndim, const float* pos, float* acc, float* pot, const float* time)
{
    if(g_nemoPot && (*ndim == 2 || *ndim == 3)) nemoAcc<float>(*ndim, *time, 1, pos, nullptr, pot, acc, 0);
}
