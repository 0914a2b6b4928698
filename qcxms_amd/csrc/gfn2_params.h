// GFN2-xTB parameters (H, C, N, O) for the HIP library.
//
// The reference takes these from tblite v0.2.1's new_gfn2_calculator
// (src/tblite.f90:123), which is not vendored with the reference.  The values
// are the published GFN2-xTB parameters (Bannwarth, Ehlert, Grimme, JCTC 15,
// 1652 (2019)) as recalled; they are not checkable against tblite here.
// Element slots: 0 = H, 1 = C, 2 = N, 3 = O.  Shell k of an element has
// angular momentum k (s, then p).
#pragma once

namespace gfn2 {

constexpr int kElem = 4;
constexpr int kMaxShellPerAtom = 2;

constexpr int kZ[kElem] = {1, 6, 7, 8};
constexpr int kNsh[kElem] = {1, 2, 2, 2};
constexpr int kPqn[kElem][2] = {{1, 0}, {2, 2}, {2, 2}, {2, 2}};  // principal quantum numbers
constexpr int kNprim[kElem][2] = {{3, 0}, {4, 4}, {4, 4}, {4, 4}};
constexpr double kSlater[kElem][2] = {
    {1.230000, 0.0}, {2.096432, 1.800000}, {2.339881, 2.014332}, {2.439742, 2.137023}};
constexpr double kSelfEnergyEv[kElem][2] = {
    {-10.707211, 0.0}, {-13.970922, -10.063292}, {-16.686243, -12.523956},
    {-20.229985, -15.503117}};
constexpr double kKcnEv[kElem][2] = {
    {-0.0500000, 0.0}, {-0.0102144, 0.0161657}, {-0.0198407, 0.0230461},
    {-0.0173888, 0.0279349}};
constexpr double kShellPoly[kElem][2] = {
    {-0.00953618, 0.0}, {-0.02294321, -0.00271102}, {-0.08506003, -0.02504201},
    {-0.14955291, -0.03350819}};
constexpr double kRefOcc[kElem][2] = {{1.0, 0.0}, {1.0, 3.0}, {1.5, 3.5}, {2.0, 4.0}};
constexpr double kHubbard[kElem] = {0.405771, 0.538015, 0.461493, 0.451896};
constexpr double kShellHubbard[kElem][2] = {
    {0.0, 0.0}, {0.0, 0.1056358}, {0.0, 0.1164892}, {0.0, 0.1497020}};
constexpr double kHubbardDeriv[kElem] = {0.0800000, 0.1500000, -0.0639780, -0.0517134};
constexpr double kRepAlpha[kElem] = {2.213717, 1.247655, 1.682689, 2.165712};
constexpr double kRepZeff[kElem] = {1.105388, 4.231078, 5.242592, 5.784415};
constexpr double kEn[kElem] = {2.20, 2.55, 3.04, 3.44};
constexpr double kAtomicRadAa[kElem] = {0.32, 0.75, 0.71, 0.64};
constexpr double kDipKernel[kElem] = {0.05563889, -0.00411674, 0.03521273, -0.04935670};
constexpr double kQuadKernel[kElem] = {0.00027431, 0.00213583, 0.02026786, -0.00310828};
constexpr double kMpRad[kElem] = {1.4, 3.0, 1.9, 1.8};
constexpr double kMpVcn[kElem] = {1.0, 3.0, 3.0, 2.0};
constexpr double kValence[kElem] = {1.0, 4.0, 5.0, 6.0};
// Pyykko & Atsumi 2009 covalent radii (Angstrom), x 4/3 for the CN
// (covalent_radii.f90:27-29, 83-84)
constexpr double kCovRad2009Aa[kElem] = {0.32, 0.75, 0.71, 0.63};

constexpr double kShellK[2] = {1.85, 2.23};  // k_s, k_p; k_sp = mean
constexpr double kEnScale = 0.02;
constexpr double kWexp = 0.5;
constexpr double kThirdShell[2] = {1.0, 0.5};
constexpr double kRepKexp = 1.5, kRepKexpLight = 1.0;
constexpr double kAesDmp3 = 3.0, kAesDmp5 = 4.0, kAesShift = 1.2, kAesKexp = 4.0,
                 kAesRmax = 5.0;
constexpr double kCnKa = 10.0, kCnKb = 20.0, kCnShift = 2.0, kCnCutoff = 25.0;
constexpr double kRepCutoff = 25.0;

// D4 (dispersion of GFN2): damping and weighting parameters
constexpr double kD4S6 = 1.0, kD4S8 = 2.7, kD4A1 = 0.52, kD4A2 = 5.0, kD4S9 = 5.0;
constexpr double kD4Alp = 16.0, kD4Wf = 6.0, kD4Ga = 3.0, kD4Gc = 2.0;
// D4 covalent CN (dftd4.f90:10-19)
constexpr double kD4K1 = 16.0, kD4K4 = 4.10451, kD4K5 = 19.08857,
                 kD4K6 = 2.0 * 11.28174 * 11.28174;
constexpr double kD4CnThr = 1600.0, kD4RThr = 5000.0, kD4AtmThr = 1600.0;

// unit conversions (xtb_mctc_convert.f90:24-63, tblite.f90:43)
constexpr double kAutoAa = 0.52917726;
constexpr double kAatoAu = 1.0 / 0.52917726;
constexpr double kAutoEv = 27.21138505;
constexpr double kEvtoAu = 1.0 / 27.21138505;
constexpr double kKtoAu = 3.166808578545117e-06;

}  // namespace gfn2
