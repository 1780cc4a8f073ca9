// Host-side construction of Multipole potentials for the agama_amd library:
//   density models -> spherical-harmonic coefficients (Poisson solver)
//   coefficient text files (the reference's format) <-> coefficients
//   coefficients -> 2-D quintic spline tables in the device layout of agx_tables.h
// All of this runs once per potential on the CPU, as in the reference; evaluation and orbit
// integration run on the GPU.
#pragma once
#include <map>
#include <memory>
#include <string>
#include <vector>
#include "agx_tables.h"

namespace agx {

// symmetry flags, same values as the reference's coord::SymmetryType (src/coord.h:50-75)
enum Sym { ST_NONE = 0, ST_XREFL = 1, ST_YREFL = 2, ST_ZREFL = 4, ST_REFL = 8, ST_ZROT = 16,
    ST_ROT = 32, ST_UNKNOWN = 64, ST_BISYM = 12, ST_TRIAXIAL = 15, ST_AXISYM = 31,
    ST_SPHERICAL = 63 };

// case-insensitive key=value parameters (the reference's INI / keyword conventions)
struct Params {
    std::map<std::string, std::string> kv;
    static Params parse(const std::string& text);   // "a=1, b=2" or one pair per line
    bool has(const std::string& key) const;
    std::string str(const std::string& key, const std::string& def = "") const;
    double num(const std::string& key, double def) const;
    void checkOnly(const std::vector<std::string>& allowed) const;   // rejects unknown keys
};

int symmetryFromName(const std::string& name);   // reference: getSymmetryTypeByName
std::string symmetryName(int sym);                // reference: getSymmetryNameByType

// spherical-harmonic coefficients on a radial grid: Phi[c][k], dPhi[c][k], c = l*(l+1)+m
struct Coefs {
    std::vector<double> radii;
    std::vector<std::vector<double> > Phi, dPhi;
    int lmax() const;
};

// one density model as plain parameters, for evaluation on the device (agx_construct.hip)
enum DensityType { DENS_DEHNEN = 1, DENS_FERRERS = 2, DENS_DISK = 3, DENS_NFW = 4, DENS_PLUMMER = 5 };
struct DensityComp {
    int type = 0;
    double p[5] = { 0, 0, 0, 0, 0 };   // Dehnen M a gamma p q; Ferrers a b c rho0; Disk Sigma0 Rd h;
                                       // NFW M rs; Plummer M a
};

struct Density {
    virtual ~Density() {}
    virtual double rhoCyl(double R, double z, double phi) const = 0;
    virtual int symmetry() const = 0;
    virtual void describe(std::vector<DensityComp>& out) const = 0;   // components, summed in order
};
std::unique_ptr<Density> makeDensity(const Params& p);   // Dehnen, Ferrers, Disk, NFW, Plummer,
                                                         // or a sum given as "a+b" components

Coefs multipoleFromDensity(const Density& dens, int lmax, int mmax, int symExp,
    int gridSizeR, double rmin, double rmax);

// The Poisson solver of multipoleFromDensity in three parts, so that its bulk can run on the
// device: the grids and the forward spherical-harmonic transform as a matrix (host), the
// density at every (quadrature radius, angular node), its transform and the radial
// integrals (device, agx_construct.hip), and the restriction to the requested order (host).
struct PoissonSetup {
    int gridSizeR = 0, lmax = 0, mmax = 0, symExp = 0, ncoef = 0, nang = 0, nint = 0, glo = 0;
    std::vector<double> R;           // grid radii [gridSizeR]
    std::vector<double> rint;        // quadrature radii [nint = (gridSizeR+1)*glo]
    std::vector<double> gln, glw;    // Gauss-Legendre nodes and weights on [0, 1] [glo]
    std::vector<double> cth, phi;    // angular nodes [nang]
    std::vector<double> B;           // forward transform [nang][ncoef]: coefs = values . B
    std::vector<int> active;         // [ncoef]: 1 for the (l, m) of the source's index set
};
PoissonSetup poissonSetup(const Density& dens, int lmax, int mmax, int symExp,
    int gridSizeR, double rmin, double rmax);
// Phi[c*gridSizeR + k], dPhi likewise, for all ncoef terms -> the coefficients of the expansion
Coefs poissonFinish(const PoissonSetup& S, const std::vector<double>& Phi, const std::vector<double>& dPhi);
// the device's middle part (density values, matrix transform with the near-zero cut, radial
// integrals) run on the host in the device's order: the CPU check of agx_construct.hip
void poissonStagesHost(const PoissonSetup& S, const Density& dens, std::vector<double>& Phi,
    std::vector<double>& dPhi);
Coefs readCoefFile(const std::string& path);
void writeCoefFile(const std::string& path, const Coefs& c);

// host copy of the device tables
struct HostPowerLaw {
    double r0sq = 0, Q = 0;
    int inner = 1, lmax = 0, mmax = 0, mmin = 0, step = 1;
    std::vector<int> lmin;          // [m + mmax]
    std::vector<double> S, U, W;    // indexed by c
};
struct HostTables {
    int nr = 0, nt = 0, ns = 0, mmax = 0, mmin = 0, lmax = 0, step = 1, logScaling = 0;
    int sym = 0;                    // symmetry of the expansion (coord::SymmetryType bits)
    double invPhi0 = 0, rsqInner = 0, rsqOuter = 0;
    std::vector<int> slotM;
    int zeroSlot = 0;
    std::vector<double> gridR, gridT, nodes;   // nodes: see agx_tables.h
    HostPowerLaw inner, outer;
    // lmax <= 2 (the reference's MultipoleInterp1d): 1-D splines per (l,m) term, no 2-D tables
    int interp1d = 0, nc1d = 0;
    int lmax1d = 0, mmax1d = 0, mmin1d = 0, step1d = 1;
    std::vector<int> lmin1d;        // [m + mmax1d]
    std::vector<double> nodes1d;    // see agx_tables.h
    std::vector<char> term1d;       // [c]: 1 for the (l,m) terms of the index set
};
// Expansions with lmax <= 2 get the reference's 1-d splines per (l,m) term (MultipoleInterp1d);
// allowLowOrder: build the 2-d tables for them instead (used internally, e.g. for the
// surrogate's l<=2 baseline)
HostTables buildTables(const Coefs& c, bool allowLowOrder = false);
// the expansion with every term of degree l > lmax set to zero
Coefs truncateCoefs(const Coefs& c, int lmax);
// Per-cell polynomial form of the 2-D tables (agx_tables.h, AGX_CELL_COEFS): for every cell of
// the (ln r, tau) grid and harmonic slot, the 36 coefficients of the bi-quintic that the
// tensor-product quintic Hermite interpolation of the cell's four corner records equals
std::vector<double> buildCellPolys(const HostTables& T);
// Bucket table of the tau grid for the FP64 kernels' cell search (agx_tables.h): a power of two
// of uniform buckets of [-1, 1], each narrower than the narrowest tau cell (at most 4096), and
// for each the index of the cell holding its left edge
std::vector<int> buildTauLookup(const HostTables& T);
// Legendre factors of the FP64 kernels' power-law extrapolation (agx_tables.h: legTab, then
// legStart appended), for degrees up to lmax
std::vector<double> buildLegendreTable(int lmax);

// the reference's sin/cos (math::sincos, src/math_core.cpp:171-206: multiples of pi/2 exact)
void sincosAgama(double x, double& s, double& c);

}  // namespace agx
