// Device-visible layout of a Multipole potential, shared by the host builder (C++) and the HIP
// kernels.  Plain structs of scalars and device pointers: passed to kernels by value.
//
// Interior (MultipoleInterp2d in the reference, src/potential_multipole.cpp:1854-2080):
//   one 2-D quintic spline in (ln r, tau) per non-empty azimuthal harmonic m ("slot").
//   All slots share the (ln r, tau) grid, so the node data are stored node-major:
//     nodes[((i*nt + j)*ns + s)*AGX_NODE_STRIDE + q],  i < nr (ln r), j < nt (tau), s < ns,
//     q = 0..8 : f, f_x, f_y, f_xx, f_xy, f_yy, f_xxy, f_xyy, f_xxyy   (q = 9: padding)
//   so a point's 4 cell corners for all harmonics are two contiguous runs of 2*ns*80 bytes.
// Per-cell polynomial form of the same splines (FP64 kernels): for cell (i, j) of the grid and
//   slot s, the 36 coefficients of the bi-quintic p(u, v) = sum_{a,b<6} c[a][b] u^a v^b,
//   u = ln r - lnr_i, v = tau - tau_j, that the tensor-product quintic Hermite interpolation of
//   the cell's corner records is:  cells[((i*(nt-1) + j)*ns + s)*AGX_CELL_COEFS + 6a + b];
//   a point's whole evaluation reads one contiguous run of ns*288 bytes.
// Interior for lmax <= 2 (MultipoleInterp1d, src/potential_multipole.cpp:1691-1815): one 1-D
//   quintic spline in ln r per (l,m) term, node-major:
//     nodes1d[(i*nc1d + c)*4 + q],  i < nr, c = l*(l+1)+m < nc1d = (lmax+1)^2,
//     q = 0..2 : f, f_x, f_xx  (q = 3: padding)
// Extrapolation (PowerLawMultipole, src/potential_multipole.cpp:1585-1658): per (l,m) term
//   S, U, W indexed by c = l*(l+1)+m, plus Q for the inner monopole.
#pragma once

#define AGX_NODE_STRIDE 10
#define AGX_CELL_COEFS 36
#define AGX_LEG_STRIDE 4
#define AGX_MAX_M 16
#define AGX_MAX_SLOTS (2 * AGX_MAX_M + 1)

struct AgxPowerLawDev {
    double r0sq;
    double Q;
    int inner;        // 1: r < rmin branch (v = l), 0: r > rmax branch (v = -l-1)
    int lmax;         // of the index set deduced from the non-zero U terms
    int mmax, mmin, step;
    int lmin[AGX_MAX_SLOTS];   // lmin(m) at [m + mmax]; > lmax means the harmonic is empty
    int uniformS;     // 1: every S of the index set equals its term's v or S[0]
    const double* S;
    const double* U;
    const double* W;
};

// index set of a spherical-harmonic expansion (math::SphHarmIndices, src/math_sphharm.h:116-157)
struct AgxSphIndDev {
    int lmax, mmax, mmin, step;
    int lmin[AGX_MAX_SLOTS];   // lmin(m) at [m + mmax]; > lmax means the harmonic is empty
};

struct AgxMultipoleDev {
    int nr, nt, ns;             // radial nodes, tau nodes, stored harmonic slots
    int mmax, mmin;             // azimuthal index range of the expansion
    int logScaling;             // m=0 term stored as ln(invPhi0 - 1/Phi), others as Phi_m/Phi_0
    int sym;                    // symmetry of the expansion (bits of the reference's coord::SymmetryType)
    double invPhi0;
    double rsqInner, rsqOuter;  // (rmin*(1+eps))^2, (rmax*(1-eps))^2 : branch thresholds
    int slotM[AGX_MAX_SLOTS];   // m of each stored slot, ascending as in the reference loop
    int zeroSlot;               // slot holding m = 0
    const double* gridR;        // ln r nodes [nr]
    const double* gridT;        // tau nodes [nt]
    const double* nodes;        // node-major spline data (see above)
    const float* nodesF;        // the same data rounded to float (precision-split form), or null
    const double* cells;        // per-cell polynomial form (see above), or null (1-D variant)
    unsigned int cellBytes;     // its size in bytes
    // cell search of the FP64 kernels: (nr-1)/(lnr_last - lnr_0) for the first guess on the
    // (near-uniform) ln r grid, and a table of nTauLut uniform buckets of [-1, 1]: tauLut[b] = the
    // tau cell holding the bucket's left edge
    double lnrScale;
    int nTauLut;
    const int* tauLut;
    // Legendre factors of the FP64 kernels' power-law extrapolation, for 0 <= m <= l <= legLmax:
    // legTab[(m*(legLmax+1) + l)*AGX_LEG_STRIDE + {0,1,2,3}] = the normalisation of P_l^m times
    // the synthesis factor (2 sqrt(pi), times sqrt(2) for m != 0), (2l-1)/(l-m) and
    // (l+m-1)/(l-m) of the recurrence in l, and l; legStart[m] = the coefficient of
    // sin^m(theta) in P_m^m before normalisation, (-1)^m (2m-1)!!
    int legLmax;
    const double* legTab;
    const double* legStart;
    AgxPowerLawDev inner, outer;
    // lmax <= 2: the 1-D variant replaces the 2-D tables (nt = ns = 0); logScaling and invPhi0
    // then refer to the l=0 coefficient Phi_00 (MultipoleInterp1d's convention)
    int interp1d;
    int nc1d;
    AgxSphIndDev ind1d;
    const double* nodes1d;
    // optional analytic NFW halo added to the expansion (a two-component Composite, as in
    // src/potential_composite.cpp:51-163): Phi = -M ln(1 + r/a) / r; nfwMass == 0: none
    double nfwMass, nfwScale;
};
